// Fused pass kernels of the Green-function matvec y = P IDFT_L(G^ . DFT_L(E x)).
//
// Every transformed axis d uses the even/odd frequency split of the length
// L_d = 2 B_d embedding (the pruned first radix-2 stage): with x[n] := 0 for
// n >= S_d,
//   even[n] = x[n] + x[n+B],  odd[n] = phi[n] (x[n] - x[n+B]),  phi[n] = e^{-i pi n/B},
// and DFT_L(E x)[2k+p] = DFT_B(p ? odd : even)[k].  The inverse merges
//   y[n] = e[n] + conj(phi[n]) o[n]      (n < min(T,B)),
//   y[n+B] = e[n] - conj(phi[n]) o[n]    (n < T-B),
// and all scale factors (1/L per axis) are folded into the stored G^.
//
// Workspace layouts (column-major, axis 1 fastest; index p*B+k on a
// transformed axis, p the parity):
//   X  (S1, S2, S3, 3)            source field
//   W1 (L1, Y1, Z, 3)             after pass 1 / before pass 5, Y1 = max(S2,T2)
//   W2 (L1, L2, Z, 3)             after pass 2 / pass 3, Z = max(S3,T3)
//   Y  (T1, T2, T3, 3)            target field
// Passes: 1 x-forward, 2 y-forward, 3 z-forward * G^ * z-inverse (in place),
// 4 y-inverse, 5 x-inverse (+ alpha/beta epilogue).
#pragma once
#include <type_traits>
#include "gla_fft.hpp"

// passes 1, 2, 4, 5: the root table and split phases of the pass's length
// read from an LDS copy made at the start of each workgroup (1) or from
// global memory (0)
#ifndef GLA_XYTAB
#define GLA_XYTAB 1
#endif
// passes 2 and 4: the two parities through the transform together (2, one
// LDS buffer each) or one after the other through one buffer (1: half the
// LDS per workgroup, so more workgroups per CU)
#ifndef GLA_YK
#define GLA_YK 2
#endif
#ifndef GLA_YK_F
#define GLA_YK_F GLA_YK
#endif
#ifndef GLA_YK_D
#define GLA_YK_D GLA_YK
#endif
template <typename R> struct YK { static constexpr int value = GLA_YK_F; };
template <> struct YK<double> { static constexpr int value = GLA_YK_D; };

namespace gla {

// Copies the length-N root table (tab[0, N)) and split phases (tab[N, 2N))
// of an axis into LDS for the workgroup's nt threads.
template <typename R, int N>
__device__ __forceinline__ void stage_tables(cpx<R>* tab, const cpx<R>* __restrict__ tw,
                                             const cpx<R>* __restrict__ phz, int nt)
{
    for (int i = threadIdx.x; i < N; i += nt) {
        tab[i] = tw[i];
        tab[N + i] = phz[i];
    }
    __syncthreads();
}

// per-axis description of how a branch frequency k maps onto stored G^
// entries (restating mulBrnHst!, src/glaActSup.jl:691-753): on a symmetric
// axis the even branch reads k for k < fe and B-k (sign -1 on off-diagonal
// components involving this axis) beyond; the odd branch reads k for k < fo
// and B-1-k beyond.
struct AxisMap {
    int B, U, sym, fe, fo;
};

__device__ __forceinline__ void axis_lookup(const AxisMap& a, int p, int k, int& j, int& s)
{
    if (!a.sym) { j = k; s = 1; return; }
    if (p == 0) {
        if (k < a.fe) { j = k; s = 1; } else { j = a.B - k; s = -1; }
    } else {
        if (k < a.fo) { j = k; s = 1; } else { j = a.B - 1 - k; s = -1; }
    }
}

// Symmetric 3x3 multiply of one branch column (mulKerHst!,
// src/glaActSup.jl:879-899): w[c][e] at axis-3 frequency k3 = t + TP e.
// gb points at the column's entry of the branch in the device layout
// gf[((b U3 + j3) U2 U1 + j2 U1 + j1) * 6 + comp], uu = U1 U2, so each point
// adds j3 * uu records; s1, s2 are the column's mirror signs.
// The six stored components of point e of a column (and its axis-3 mirror
// sign), for the pipelined multiply below.
template <typename R, int TP>
__device__ __forceinline__ void g_fetch(const cpx<R>* __restrict__ gb, int uu, const AxisMap& a3, int p3, int t, int e,
                                        cpx<R> (&d)[6], R& s3f)
{
    int j3, s3;
    axis_lookup(a3, p3, t + TP * e, j3, s3);
    s3f = R(s3);
    const cpx<R>* g = gb + (int64_t)(j3 * uu) * 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = g[k];
}

// PIPE: the six loads of point e + 1 are issued before point e is
// multiplied, so one memory latency is exposed per call instead of one per
// point; with `first` the caller has already fetched point 0 (g_fetch), for
// instance before the forward transform, whose stages then hide its latency.
template <typename R, int E, int TP, bool PIPE = false>
__device__ __forceinline__ void g_multiply(cpx<R> (&w)[3][E], const cpx<R>* __restrict__ gb, int uu,
                                           const AxisMap& a3, int p3, int t, int s1, int s2, bool act,
                                           const cpx<R> (*first)[6] = nullptr, R first_s3 = R(1))
{
    const R s12 = R(s1 * s2);
    if constexpr (PIPE) {
        cpx<R> gv[2][6];
        R s3v[2];
        auto fetch = [&](int e, cpx<R>(&d)[6], R& s3f) { g_fetch<R, TP>(gb, uu, a3, p3, t, e, d, s3f); };
        if (!act) return;
        if (first) {
#pragma unroll
            for (int k = 0; k < 6; ++k) gv[0][k] = (*first)[k];
            s3v[0] = first_s3;
        } else {
            fetch(0, gv[0], s3v[0]);
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e + 1 < E) fetch(e + 1, gv[(e + 1) & 1], s3v[(e + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
            const cpx<R>* g = gv[e & 1];
            const R s13 = R(s1) * s3v[e & 1], s23 = R(s2) * s3v[e & 1];
            const cpx<R> gxy = s12 * g[3], gxz = s13 * g[4], gyz = s23 * g[5];
            const cpx<R> x0 = w[0][e], x1 = w[1][e], x2 = w[2][e];
            w[0][e] = g[0] * x0 + gxy * x1 + gxz * x2;
            w[1][e] = gxy * x0 + g[1] * x1 + gyz * x2;
            w[2][e] = gxz * x0 + gyz * x1 + g[2] * x2;
            __builtin_amdgcn_sched_barrier(0);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k3 = t + TP * e;
        int j3, s3;
        axis_lookup(a3, p3, k3, j3, s3);
        if (act) {
            const cpx<R>* g = gb + (int64_t)(j3 * uu) * 6;
            const cpx<R> gxx = g[0], gyy = g[1], gzz = g[2];
            const R s13 = R(s1 * s3), s23 = R(s2 * s3);
            const cpx<R> gxy = s12 * g[3], gxz = s13 * g[4], gyz = s23 * g[5];
            const cpx<R> x0 = w[0][e], x1 = w[1][e], x2 = w[2][e];
            w[0][e] = gxx * x0 + gxy * x1 + gxz * x2;
            w[1][e] = gxy * x0 + gyy * x1 + gyz * x2;
            w[2][e] = gxz * x0 + gyz * x1 + gzz * x2;
        }
    }
}

// --------------------------------------------------------------------------
// Pass 1: x-forward on contiguous rows.  One pencil = one (i2,i3,c) row of X.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PEN, int MINW>
__global__ __launch_bounds__(TP * PEN, MINW) void k_fwd_x(const cpx<R>* __restrict__ X, cpx<R>* __restrict__ W1,
                                                          int S1, int S2, int S3, int Y1, int Z,
                                                          int64_t xs, int64_t ws,
                                                          const cpx<R>* __restrict__ phz,
                                                          const cpx<R>* __restrict__ tw)
{
    constexpr int E = N / TP;
    // one padding slot every 8 positions (stride-8 stage writes hit distinct banks)
    constexpr int PADSH = 3;
    constexpr int PS = N + (N >> PADSH) + 1;
    // rows are whole pencils inside one wave when TP divides 64
    constexpr int XSYNC = (64 % TP == 0) ? 1 : 0;
    __shared__ cpx<R> lds[2 * PEN * PS];
    __shared__ cpx<R> tab[GLA_XYTAB ? 2 * N : 1];
    if constexpr (GLA_XYTAB != 0) {
        stage_tables<R, N>(tab, tw, phz, TP * PEN);
        tw = tab;
        phz = tab + N;
    }
    const int t = threadIdx.x % TP, pen = threadIdx.x / TP;
    const int rows = S2 * S3 * 3;
    const int row = blockIdx.x * PEN + pen;
    const bool act = row < rows;
    const int i2 = act ? row % S2 : 0, i3 = act ? (row / S2) % S3 : 0, c = act ? row / (S2 * S3) : 0;
    const cpx<R>* xr = X + (int64_t)blockIdx.y * xs + (int64_t)row * S1;
    cpx<R>* wr = W1 + (int64_t)blockIdx.y * ws + (((int64_t)c * Z + i3) * Y1 + i2) * (2 * N);
    cpx<R> v[2][E];   // even, odd parity
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
        if (act && n < S1) a = xr[n];
        if (act && n + N < S1) b = xr[n + N];
        v[0][e] = a + b;
        v[1][e] = phz[n] * (a - b);
    }
    PencilFft<R, N, TP, 1, PS, false, XSYNC, 2, PEN * PS, PADSH>::run(v, lds, t, pen, tw);
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            wr[t + TP * e] = v[0][e];
            wr[N + t + TP * e] = v[1][e];
        }
    }
}

// --------------------------------------------------------------------------
// Passes 1 and 2 fused (B1 = B2 = N): one workgroup per (plane (i3, c), x
// parity p1, right-hand side).  Phase 1 transforms the plane's rows (x split
// for parity p1, length-N FFT) into an LDS tile, row i2 at tile[i2*RS ..];
// each row pencil uses its own tile row as exchange space.  Phase 2 reads the
// tile's columns (y split into both parities, two length-N FFTs, each column
// pencil using its own column slots as exchange space) and writes W2
// directly: the intermediate W1 of the two-pass form never goes to HBM.
// Rows beyond S2 are scratch for masked pencils; the tile holds
// XY_ROWS(N) >= max(S2 rounded to the row-pencil count, N) rows.
// --------------------------------------------------------------------------
template <int N>
struct XyTile {
    static constexpr int RS = (N + (N >> 3) + 1) | 1;   // odd pitch: column accesses spread over banks
    static constexpr int SLOTS = 20480;                  // 160 KB of ComplexF32
    static constexpr int ROWS = SLOTS / RS;
};

template <typename R, int N, int TP1, int TP2, int NT>
__global__ __launch_bounds__(NT, 1) void k_fwd_xy(const cpx<R>* __restrict__ X, cpx<R>* __restrict__ W2, int S1,
                                                  int S2, int S3, int L1, int Z, int64_t xs, int64_t w2s,
                                                  const cpx<R>* __restrict__ phz1, const cpx<R>* __restrict__ tw1,
                                                  const cpx<R>* __restrict__ phz2, const cpx<R>* __restrict__ tw2)
{
    using T = XyTile<N>;
    constexpr int RS = T::RS;
    constexpr int E1 = N / TP1, E2 = N / TP2;
    constexpr int PEN1 = NT / TP1, CPB = NT / TP2;
    static_assert(64 % TP1 == 0 && 64 % TP2 == 0 && NT % 64 == 0, "pencils inside one wave");
    static_assert(N % CPB == 0, "whole column groups");
    static_assert(T::ROWS >= N, "the tile holds N rows");
    __shared__ cpx<R> tile[T::SLOTS];
    const int pl = blockIdx.x;
    const int c = pl / S3, i3 = pl - c * S3;
    const int p1 = blockIdx.y;
    const int64_t rhs = blockIdx.z;

    // phase 1: rows i2 of the plane, x split for parity p1 and x-FFT
    {
        const int t = threadIdx.x % TP1, pen = threadIdx.x / TP1;
        const cpx<R>* xp = X + rhs * xs + ((int64_t)c * S3 + i3) * S2 * (int64_t)S1;
        for (int r0 = 0; r0 < S2; r0 += PEN1) {
            const int i2 = r0 + pen;
            const bool act = i2 < S2;
            const cpx<R>* xr = xp + (int64_t)(act ? i2 : 0) * S1;
            cpx<R> v[1][E1];
#pragma unroll
            for (int e = 0; e < E1; ++e) {
                const int n = t + TP1 * e;
                cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
                if (act && n < S1) a = xr[n];
                if (act && n + N < S1) b = xr[n + N];
                v[0][e] = p1 == 0 ? a + b : phz1[n] * (a - b);
            }
            PencilFft<R, N, TP1, 1, RS, false, 1, 1, 0, 3>::run(v, tile, t, i2, tw1);
#pragma unroll
            for (int e = 0; e < E1; ++e) tile[i2 * RS + t + TP1 * e] = v[0][e];
        }
    }
    __syncthreads();

    // phase 2: columns k1 of the tile, y split into both parities and y-FFTs
    {
        const int t = threadIdx.x % TP2, pen = threadIdx.x / TP2;
        cpx<R>* dst = W2 + rhs * w2s + ((int64_t)c * Z + i3) * (int64_t)(2 * N) * L1 + (int64_t)p1 * N;
        for (int k0 = 0; k0 < N; k0 += CPB) {
            const int k1 = k0 + pen;
            cpx<R> a[E2], b[E2];
#pragma unroll
            for (int e = 0; e < E2; ++e) {
                const int n = t + TP2 * e;
                a[e] = n < S2 ? tile[n * RS + k1] : cpx<R>{R(0), R(0)};
                b[e] = n + N < S2 ? tile[(n + N) * RS + k1] : cpx<R>{R(0), R(0)};
            }
            lds_sync<1>();   // the column's reads are done before its slots become exchange space
#pragma unroll
            for (int p2 = 0; p2 < 2; ++p2) {
                cpx<R> v[1][E2];
#pragma unroll
                for (int e = 0; e < E2; ++e) {
                    const int n = t + TP2 * e;
                    v[0][e] = p2 == 0 ? a[e] + b[e] : phz2[n] * (a[e] - b[e]);
                }
                PencilFft<R, N, TP2, RS, 1, false, 1, 1, 0, 0>::run(v, tile, t, k1, tw2);
#pragma unroll
                for (int e = 0; e < E2; ++e)
                    dst[(int64_t)(p2 * N + t + TP2 * e) * L1 + k1] = v[0][e];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Pass 5: x-inverse + merge/crop onto T1, optional y = alpha*res + beta*y.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PEN, int MINW>
__global__ __launch_bounds__(TP * PEN, MINW) void k_inv_x(const cpx<R>* __restrict__ W1, cpx<R>* __restrict__ Y,
                                                          int T1, int T2, int T3, int Y1, int Z,
                                                          int64_t ws, int64_t ys,
                                                          const cpx<R>* __restrict__ phz,
                                                          const cpx<R>* __restrict__ tw,
                                                          cpx<R> alpha, cpx<R> beta, int axpby,
                                                          const cpx<R>* __restrict__ lsx, int64_t lsxs,
                                                          const cpx<R>* __restrict__ chi, int64_t chis)
{
    // axpby 0: y = G x;  1: y = alpha G x + beta y;  2: y = x - chi (G x),
    // chi one value per target cell (per right-hand side when chis > 0)
    constexpr int E = N / TP;
    constexpr int PADSH = 3;
    constexpr int PS = N + (N >> PADSH) + 1;
    constexpr int XSYNC = (64 % TP == 0) ? 1 : 0;
    __shared__ cpx<R> lds[2 * PEN * PS];
    __shared__ cpx<R> tab[GLA_XYTAB ? 2 * N : 1];
    if constexpr (GLA_XYTAB != 0) {
        stage_tables<R, N>(tab, tw, phz, TP * PEN);
        tw = tab;
        phz = tab + N;
    }
    const int t = threadIdx.x % TP, pen = threadIdx.x / TP;
    const int rows = T2 * T3 * 3;
    const int row = blockIdx.x * PEN + pen;
    const bool act = row < rows;
    const int i2 = act ? row % T2 : 0, i3 = act ? (row / T2) % T3 : 0, c = act ? row / (T2 * T3) : 0;
    const cpx<R>* wr = W1 + (int64_t)blockIdx.y * ws + (((int64_t)c * Z + i3) * Y1 + i2) * (2 * N);
    cpx<R>* yr = Y + (int64_t)blockIdx.y * ys + (int64_t)row * T1;
    const int64_t cell0 = ((int64_t)i3 * T2 + i2) * T1;
    cpx<R> v[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TP * e;
        v[0][e] = act ? wr[k] : cpx<R>{R(0), R(0)};
        v[1][e] = act ? wr[N + k] : cpx<R>{R(0), R(0)};
    }
    PencilFft<R, N, TP, 1, PS, true, XSYNC, 2, PEN * PS, PADSH>::run(v, lds, t, pen, tw);
    if (!act) return;
    const int top = T1 < N ? T1 : N;
    const bool rd = axpby == 1 && !(beta.x == R(0) && beta.y == R(0));
    const cpx<R>* xr = axpby == 2 ? lsx + (int64_t)blockIdx.y * lsxs + (int64_t)row * T1 : nullptr;
    const cpx<R>* cr = axpby == 2 ? chi + (int64_t)blockIdx.y * chis + cell0 : nullptr;
    auto epilogue = [&](cpx<R> r, int n) {
        if (axpby == 1) r = alpha * r + (rd ? beta * yr[n] : cpx<R>{R(0), R(0)});
        else if (axpby == 2) r = xr[n] - cr[n] * r;
        yr[n] = r;
    };
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        const cpx<R> po = cmulc(v[1][e], phz[n]);
        if (n < top) epilogue(v[0][e] + po, n);
        if (n + N < T1) epilogue(v[0][e] - po, n + N);
    }
}

// --------------------------------------------------------------------------
// Pass 2: y-forward on bundles of PX consecutive axis-1 columns.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PX, int MINW>
__global__ __launch_bounds__(TP * PX, MINW) void k_fwd_y(const cpx<R>* __restrict__ W1, cpx<R>* __restrict__ W2,
                                                   int L1, int S2, int S3, int Y1, int Z, int64_t w1s,
                                                   int64_t w2s, const cpx<R>* __restrict__ phz,
                                                   const cpx<R>* __restrict__ tw)
{
    constexpr int E = N / TP;
    __shared__ cpx<R> lds[YK<R>::value * PX * N];
    __shared__ cpx<R> tab[GLA_XYTAB ? 2 * N : 1];
    if constexpr (GLA_XYTAB != 0) {
        stage_tables<R, N>(tab, tw, phz, TP * PX);
        tw = tab;
        phz = tab + N;
    }
    const int xl = threadIdx.x % PX, t = threadIdx.x / PX;
    const int m1 = blockIdx.x * PX + xl;
    const bool act = m1 < L1;
    const int i3 = blockIdx.y % S3, c = blockIdx.y / S3;
    const cpx<R>* src = W1 + (int64_t)blockIdx.z * w1s + ((int64_t)c * Z + i3) * Y1 * L1 + m1;
    cpx<R>* dst = W2 + (int64_t)blockIdx.z * w2s + ((int64_t)c * Z + i3) * (int64_t)(2 * N) * L1 + m1;
    cpx<R> v[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
        if (act && n < S2) a = src[(int64_t)n * L1];
        if (act && n + N < S2) b = src[(int64_t)(n + N) * L1];
        v[0][e] = a + b;
        v[1][e] = phz[n] * (a - b);
    }
    if constexpr (YK<R>::value == 2) {
        PencilFft<R, N, TP, PX, 1, false, 0, 2, PX * N>::run(v, lds, t, xl, tw);
    } else {
        PencilFft<R, N, TP, PX, 1, false>::run(v[0], lds, t, xl, tw);
        PencilFft<R, N, TP, PX, 1, false>::run(v[1], lds, t, xl, tw);
    }
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            dst[(int64_t)(t + TP * e) * L1] = v[0][e];
            dst[(int64_t)(N + t + TP * e) * L1] = v[1][e];
        }
    }
}

// --------------------------------------------------------------------------
// Pass 4: y-inverse + merge/crop onto T2.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PX, int MINW>
__global__ __launch_bounds__(TP * PX, MINW) void k_inv_y(const cpx<R>* __restrict__ W2, cpx<R>* __restrict__ W1,
                                                   int L1, int T2, int T3, int Y1, int Z, int64_t w2s,
                                                   int64_t w1s, const cpx<R>* __restrict__ phz,
                                                   const cpx<R>* __restrict__ tw)
{
    constexpr int E = N / TP;
    __shared__ cpx<R> lds[YK<R>::value * PX * N];
    __shared__ cpx<R> tab[GLA_XYTAB ? 2 * N : 1];
    if constexpr (GLA_XYTAB != 0) {
        stage_tables<R, N>(tab, tw, phz, TP * PX);
        tw = tab;
        phz = tab + N;
    }
    const int xl = threadIdx.x % PX, t = threadIdx.x / PX;
    const int m1 = blockIdx.x * PX + xl;
    const bool act = m1 < L1;
    const int i3 = blockIdx.y % T3, c = blockIdx.y / T3;
    const cpx<R>* src = W2 + (int64_t)blockIdx.z * w2s + ((int64_t)c * Z + i3) * (int64_t)(2 * N) * L1 + m1;
    cpx<R>* dst = W1 + (int64_t)blockIdx.z * w1s + ((int64_t)c * Z + i3) * Y1 * L1 + m1;
    cpx<R> v[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TP * e;
        v[0][e] = act ? src[(int64_t)k * L1] : cpx<R>{R(0), R(0)};
        v[1][e] = act ? src[(int64_t)(N + k) * L1] : cpx<R>{R(0), R(0)};
    }
    if constexpr (YK<R>::value == 2) {
        PencilFft<R, N, TP, PX, 1, true, 0, 2, PX * N>::run(v, lds, t, xl, tw);
    } else {
        PencilFft<R, N, TP, PX, 1, true>::run(v[0], lds, t, xl, tw);
        PencilFft<R, N, TP, PX, 1, true>::run(v[1], lds, t, xl, tw);
    }
    if (!act) return;
    const int top = T2 < N ? T2 : N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        const cpx<R> po = cmulc(v[1][e], phz[n]);
        if (n < top) dst[(int64_t)n * L1] = v[0][e] + po;
        if (n + N < T2) dst[(int64_t)(n + N) * L1] = v[0][e] - po;
    }
}

// --------------------------------------------------------------------------
// Pass 3: z-forward, G^ multiply, z-inverse, merge/crop onto T3 -- in place
// on W2 (each workgroup owns its columns: every load precedes the first
// barrier after which any store can happen, every store follows the last
// load).  Each thread carries all three
// field components; KB = 3 runs the three transforms through each stage
// together (three LDS buffers), KB = 1 one after the other (one buffer).
// MODE 0: S3 == T3 == B3; MODE 1: S3 > B3 (spill-over add); MODE 2: T3 > B3.
// G^ device layout: gf[(((b*U3 + j3)*U2 + j2)*U1 + j1)*6 + comp],
// comp order [xx,yy,zz,xy,xz,yz], pre-scaled by 1/(L1 L2 L3).
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PX, int MODE, int MINW, int KB, int RELOAD, int WL, int GG, int PADSH>
__global__ __launch_bounds__(TP * PX, MINW) void k_zpass(cpx<R>* __restrict__ W2, int L1, int L2, int S3, int T3,
                                                   int Z, int64_t w2s, const cpx<R>* __restrict__ gf,
                                                   AxisMap a1, AxisMap a2, AxisMap a3,
                                                   const cpx<R>* __restrict__ phz,
                                                   const cpx<R>* __restrict__ tw,
                                                   const int* __restrict__ order, int per_xcd, int nrhs)
{
    constexpr int E = N / TP;
    // WL: every pencil lives inside one wave (64/TP columns per wave, each
    // column's TP threads adjacent), in its own padded LDS row, so the
    // exchanges between stages need no workgroup barrier
    static_assert(!WL || (64 % TP == 0 && (TP * PX) % 64 == 0), "wave-local pencils need TP | 64");
    constexpr int CPW = WL ? 64 / TP : 1;
    constexpr int PSW = WL ? N + (N >> PADSH) + 1 : N;
    __shared__ cpx<R> lds[KB * PX * PSW];
    int xl, t;
    if constexpr (WL) {
        const int lane = threadIdx.x & 63;
        xl = (threadIdx.x >> 6) * CPW + lane % CPW;
        t = lane / CPW;
    } else {
        xl = threadIdx.x % PX;
        t = threadIdx.x / PX;
    }
    // stride-interleaved pencils (block barrier) or padded pencil rows (wave sync)
    using FftF3 = std::conditional_t<WL, PencilFft<R, N, TP, 1, PSW, false, 1, 3, PX * PSW, PADSH>,
                                     PencilFft<R, N, TP, PX, 1, false, 0, 3, PX * N>>;
    using FftI3 = std::conditional_t<WL, PencilFft<R, N, TP, 1, PSW, true, 1, 3, PX * PSW, PADSH>,
                                     PencilFft<R, N, TP, PX, 1, true, 0, 3, PX * N>>;
    using FftF1 = std::conditional_t<WL, PencilFft<R, N, TP, 1, PSW, false, 1, 1, 0, PADSH>,
                                     PencilFft<R, N, TP, PX, 1, false>>;
    using FftI1 = std::conditional_t<WL, PencilFft<R, N, TP, 1, PSW, true, 1, 1, 0, PADSH>,
                                     PencilFft<R, N, TP, PX, 1, true>>;
    // workgroup h runs the column tile order[l]: the table places tiles that
    // share stored G^ entries (mirror images on symmetric axes) next to each
    // other in l, and l is chosen so that consecutive l run on one XCD
    // (workgroups are dealt round-robin over the 8 XCDs)
    // With several right-hand sides the right-hand side is the fastest index
    // of an XCD's stretch (grid 8 * per_xcd * nrhs): consecutive workgroups
    // on one XCD run the same tile for every right-hand side, so its G^
    // entries come from HBM once per launch rather than once per right-hand
    // side.
    const int h = blockIdx.x;
    const int q = h / 8;
    const int l = (h % 8) * per_xcd + q / nrhs;
    const int rhs = q % nrhs;
    const int tile = order[l];
    if (tile < 0) return;   // padding entry (whole workgroup)
    const int m1 = (tile & 0xffff) * PX + xl;
    const int m2 = tile >> 16;
    const bool act = m1 < L1;
    const int64_t cs = (int64_t)Z * L2 * L1;      // component stride
    cpx<R>* col = W2 + (int64_t)rhs * w2s + (int64_t)m2 * L1 + m1;
    const int64_t zs = (int64_t)L2 * L1;

    // RELOAD: read the column again for the odd parity (an L2 hit) instead of
    // keeping it in registers through the even parity's transforms
    constexpr int XE = RELOAD ? 1 : E;
    cpx<R> xa[3][XE];
    cpx<R> xb[MODE == 1 ? 3 : 1][MODE == 1 ? XE : 1];
    if constexpr (RELOAD == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                xa[c][e] = (act && n < S3) ? col[c * cs + (int64_t)n * zs] : cpx<R>{R(0), R(0)};
                if constexpr (MODE == 1)
                    xb[c][e] = (act && n + N < S3) ? col[c * cs + (int64_t)(n + N) * zs] : cpx<R>{R(0), R(0)};
            }
    }

    // per-column branch parities and stored indices on axes 1 and 2
    const int p1 = m1 >= a1.B ? 1 : 0, k1 = m1 - p1 * a1.B;
    const int p2 = m2 >= a2.B ? 1 : 0, k2 = m2 - p2 * a2.B;
    int j1, s1, j2, s2;
    axis_lookup(a1, p1, act ? k1 : 0, j1, s1);
    axis_lookup(a2, p2, k2, j2, s2);

    const int uu = a1.U * a2.U;
    const int gcol = j2 * a1.U + j1;
    // GG bit 1: G^ loads pipelined one point ahead; bit 2 in addition: the
    // first point's loads issued before the forward transform (pre_g), so
    // they arrive during its stages
    constexpr bool GPIPE = (GG & 2) != 0, GPRE = GPIPE && (GG & 4) != 0;
    cpx<R> g0[6];
    R g0s = R(1);
    auto gbase = [&](int p3) {
        const int b = (p1 << 2) | (p2 << 1) | p3;
        return gf + ((int64_t)b * a3.U * uu + gcol) * 6;
    };
    auto pre_g = [&](int p3) {
        if constexpr (GPRE)
            if (act) g_fetch<R, TP>(gbase(p3), uu, a3, p3, t, 0, g0, g0s);
    };
    auto mul_g = [&](cpx<R>(&w)[3][E], int p3) {
        g_multiply<R, E, TP, GPIPE>(w, gbase(p3), uu, a3, p3, t, s1, s2, act, GPRE ? &g0 : nullptr, g0s);
    };
    // GG bit 0: the transform's root table and the split phases staged in
    // LDS once per workgroup (their reads inside the stages and the split /
    // merge then wait on LDS, not on the L2)
    __shared__ cpx<R> twl[(GG & 1) ? 2 * N : 1];
    if constexpr ((GG & 1) != 0) {
        for (int i = threadIdx.x; i < N; i += TP * PX) {
            twl[i] = tw[i];
            twl[N + i] = phz[i];
        }
        __syncthreads();
    }
    const cpx<R>* twu = (GG & 1) ? twl : tw;
    const cpx<R>* phu = (GG & 1) ? twl + N : phz;
    auto fwd = [&](cpx<R>(&w)[3][E]) {
        if constexpr (KB == 3) {
            FftF3::run(w, lds, t, xl, twu);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) FftF1::run(w[c], lds, t, xl, twu);
        }
    };
    auto inv = [&](cpx<R>(&w)[3][E]) {
        if constexpr (KB == 3) {
            FftI3::run(w, lds, t, xl, twu);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) FftI1::run(w[c], lds, t, xl, twu);
        }
    };

    if constexpr (RELOAD == 2) {
        // both parities split first: the input dies at the split, so the
        // live state is the two parities' fields only
        cpx<R> wv[2][3][E];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                const cpx<R> a = (act && n < S3) ? col[c * cs + (int64_t)n * zs] : cpx<R>{R(0), R(0)};
                cpx<R> b = {R(0), R(0)};
                if constexpr (MODE == 1)
                    b = (act && n + N < S3) ? col[c * cs + (int64_t)(n + N) * zs] : cpx<R>{R(0), R(0)};
                wv[0][c][e] = a + b;
                wv[1][c][e] = phu[n] * (a - b);
            }
        pre_g(0);
        fwd(wv[0]);
        mul_g(wv[0], 0);
        inv(wv[0]);
        pre_g(1);
        fwd(wv[1]);
        mul_g(wv[1], 1);
        inv(wv[1]);
        if (!act) return;
        const int top = T3 < N ? T3 : N;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                const cpx<R> po = cmulc(wv[1][c][e], phu[n]);
                if (n < top) col[c * cs + (int64_t)n * zs] = wv[0][c][e] + po;
                if constexpr (MODE == 2)
                    if (n + N < T3) col[c * cs + (int64_t)(n + N) * zs] = wv[0][c][e] - po;
            }
        return;
    }

    cpx<R> acc[3][E];
    cpx<R> accb[MODE == 2 ? 3 : 1][MODE == 2 ? E : 1];
#pragma unroll
    for (int p3 = 0; p3 < 2; ++p3) {
        cpx<R> w[3][E];
        if constexpr (RELOAD) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                cpx<R> a, b = {R(0), R(0)};
                if constexpr (RELOAD) {
                    a = (act && n < S3) ? col[c * cs + (int64_t)n * zs] : cpx<R>{R(0), R(0)};
                    if constexpr (MODE == 1)
                        b = (act && n + N < S3) ? col[c * cs + (int64_t)(n + N) * zs] : cpx<R>{R(0), R(0)};
                } else {
                    a = xa[c][e];
                    if constexpr (MODE == 1) b = xb[c][e];
                }
                w[c][e] = p3 == 0 ? a + b : phu[n] * (a - b);
            }
        pre_g(p3);
        fwd(w);
        mul_g(w, p3);
        inv(w);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (p3 == 0) {
                    acc[c][e] = w[c][e];
                    if constexpr (MODE == 2) accb[c][e] = w[c][e];
                } else {
                    const int n = t + TP * e;
                    const cpx<R> po = cmulc(w[c][e], phu[n]);
                    acc[c][e] = acc[c][e] + po;
                    if constexpr (MODE == 2) accb[c][e] = accb[c][e] - po;
                }
            }
    }
    if (!act) return;
    const int top = T3 < N ? T3 : N;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + TP * e;
            if (n < top) col[c * cs + (int64_t)n * zs] = acc[c][e];
            if constexpr (MODE == 2)
                if (n + N < T3) col[c * cs + (int64_t)(n + N) * zs] = accb[c][e];
        }
}

// --------------------------------------------------------------------------
// Pass 3, software-pipelined: a persistent form of k_zpass for S3 <= B3.
// Each workgroup walks its XCD's part of the tile order (tiles l = xcd *
// per_xcd + q + i * bpx for its slot q of bpx per XCD) and loads the next
// tile's columns while the current one is transformed, so the memory phase of
// one tile overlaps the arithmetic of the other.  The next tile's loads are
// issued after the even parity's G^ multiply: vector-memory completions are
// counted in order, so issuing them earlier would make the even parity's G^
// loads wait for them.  Same arithmetic as k_zpass with both parities split
// first.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PX, int MODE, int KB, int PF>
__global__ __launch_bounds__(TP * PX, 2) void k_zpipe(cpx<R>* __restrict__ W2, int L1, int L2, int S3, int T3,
                                                   int Z, int64_t w2s, const cpx<R>* __restrict__ gf,
                                                   AxisMap a1, AxisMap a2, AxisMap a3,
                                                   const cpx<R>* __restrict__ phz,
                                                   const cpx<R>* __restrict__ tw,
                                                   const int* __restrict__ order, int per_xcd, int nrhs)
{
    static_assert(MODE != 1, "the pipelined pass needs S3 <= B3");
    (void)nrhs;   // right-hand side from blockIdx.z in this form
    constexpr int E = N / TP;
    __shared__ cpx<R> lds[KB * PX * N];
    int xl = threadIdx.x % PX;
    int t = threadIdx.x / PX;
    const int xcd = blockIdx.x % 8, bpx = gridDim.x / 8;
    int il = blockIdx.x / 8;
    if (il >= per_xcd) return;   // whole workgroup
    const int64_t cs = (int64_t)Z * L2 * L1;
    const int64_t zs = (int64_t)L2 * L1;
    cpx<R>* const base = W2 + (int64_t)blockIdx.z * w2s;
    const int uu = a1.U * a2.U;

    auto fwd = [&](cpx<R>(&w)[3][E]) {
        if constexpr (KB == 3) {
            PencilFft<R, N, TP, PX, 1, false, 0, 3, PX * N>::run(w, lds, t, xl, tw);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) PencilFft<R, N, TP, PX, 1, false>::run(w[c], lds, t, xl, tw);
        }
    };
    auto inv = [&](cpx<R>(&w)[3][E]) {
        if constexpr (KB == 3) {
            PencilFft<R, N, TP, PX, 1, true, 0, 3, PX * N>::run(w, lds, t, xl, tw);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) PencilFft<R, N, TP, PX, 1, true>::run(w[c], lds, t, xl, tw);
        }
    };
    // this thread's column of a tile (order entry tile | m2 << 16, -1 padding)
    auto column = [&](int tile, int& m1, int& m2) {
        m1 = (tile & 0xffff) * PX + xl;
        m2 = tile >> 16;
        return tile >= 0 && m1 < L1;
    };
    // components [c0, c1) of a tile's columns
    auto load = [&](int tile, cpx<R>(&xa)[3][E], int c0, int c1) {
        int m1, m2;
        const bool ok = column(tile, m1, m2);
        const cpx<R>* col = base + (ok ? (int64_t)m2 * L1 + m1 : 0);
#pragma unroll
        for (int c = c0; c < c1; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                xa[c][e] = (ok && n < S3) ? col[c * cs + (int64_t)n * zs] : cpx<R>{R(0), R(0)};
            }
    };

    int tile = order[xcd * per_xcd + il];
    cpx<R> xa[3][E];
    load(tile, xa, 0, 3);
    for (;;) {
        // table pointers opaque per tile: keeps the twiddle and phase loads
        // inside the loop instead of holding their values in registers
        // across all tiles
        // the thread index opaque per tile: what derives from it (offsets,
        // phases, mirror lookups) is recomputed per tile rather than held in
        // registers across the loop
        asm volatile("" : "+v"(t), "+v"(xl));
        int m1, m2;
        const bool act = column(tile, m1, m2);
        const int p1 = (act && m1 >= a1.B) ? 1 : 0, k1 = act ? m1 - p1 * a1.B : 0;
        const int p2 = (act && m2 >= a2.B) ? 1 : 0, k2 = act ? m2 - p2 * a2.B : 0;
        int j1, s1, j2, s2;
        axis_lookup(a1, p1, k1, j1, s1);
        axis_lookup(a2, p2, k2, j2, s2);
        const cpx<R>* gc = gf + (int64_t)(j2 * a1.U + j1) * 6;
        cpx<R> wv[2][3][E];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                wv[0][c][e] = xa[c][e];
                wv[1][c][e] = phz[n] * xa[c][e];
            }
        const int nil = il + bpx;
        const bool more = nil < per_xcd;   // the same for the whole workgroup
        const int ntile = more ? order[xcd * per_xcd + nil] : -1;
        const int b0 = (p1 << 2) | (p2 << 1);
        fwd(wv[0]);
        g_multiply<R, E, TP>(wv[0], gc + (int64_t)b0 * a3.U * uu * 6, uu, a3, 0, t, s1, s2, act);
        // the next tile's first PF components load during the rest of this
        // tile's transforms, the others during the odd parity's inverse
        __builtin_amdgcn_sched_barrier(0);
        if (more && PF > 0) load(ntile, xa, 0, PF);
        __builtin_amdgcn_sched_barrier(0);
        inv(wv[0]);
        fwd(wv[1]);
        g_multiply<R, E, TP>(wv[1], gc + (int64_t)(b0 | 1) * a3.U * uu * 6, uu, a3, 1, t, s1, s2, act);
        if constexpr (PF >= 0 && PF < 3) {
            __builtin_amdgcn_sched_barrier(0);
            if (more) load(ntile, xa, PF, 3);
            __builtin_amdgcn_sched_barrier(0);
        }
        inv(wv[1]);
        if (act) {
            cpx<R>* col = base + (int64_t)m2 * L1 + m1;
            const int top = T3 < N ? T3 : N;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int n = t + TP * e;
                    const cpx<R> po = cmulc(wv[1][c][e], phz[n]);
                    if (n < top) col[c * cs + (int64_t)n * zs] = wv[0][c][e] + po;
                    if constexpr (MODE == 2)
                        if (n + N < T3) col[c * cs + (int64_t)(n + N) * zs] = wv[0][c][e] - po;
                }
        }
        if (!more) break;
        il = nil;
        tile = ntile;
        if constexpr (PF < 0) load(tile, xa, 0, 3);
    }
}

// --------------------------------------------------------------------------
// Forward z pass without the multiply (operator setup): split S3 -> 2 x B3
// and transform, in place on W2 (L1, L2, Z, 3) per right-hand side.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PX, int MINW>
__global__ __launch_bounds__(TP * PX, MINW) void k_fwd_z(cpx<R>* __restrict__ W2, int L1, int L2, int S3, int Z,
                                                   int64_t w2s, const cpx<R>* __restrict__ phz,
                                                   const cpx<R>* __restrict__ tw)
{
    constexpr int E = N / TP;
    __shared__ cpx<R> lds[2 * PX * N];
    const int xl = threadIdx.x % PX, t = threadIdx.x / PX;
    const int m1 = blockIdx.x * PX + xl;
    const bool act = m1 < L1;
    const int m2 = blockIdx.y % L2, c = blockIdx.y / L2;
    const int64_t zs = (int64_t)L2 * L1;
    cpx<R>* col = W2 + (int64_t)blockIdx.z * w2s + (int64_t)c * Z * zs + (int64_t)m2 * L1 + m1;
    cpx<R> v[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
        if (act && n < S3) a = col[(int64_t)n * zs];
        if (act && n + N < S3) b = col[(int64_t)(n + N) * zs];
        v[0][e] = a + b;
        v[1][e] = phz[n] * (a - b);
    }
    PencilFft<R, N, TP, PX, 1, false, 0, 2, PX * N>::run(v, lds, t, xl, tw);
    if (!act) return;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        col[(int64_t)(t + TP * e) * zs] = v[0][e];
        col[(int64_t)(N + t + TP * e) * zs] = v[1][e];
    }
}

}  // namespace gla
