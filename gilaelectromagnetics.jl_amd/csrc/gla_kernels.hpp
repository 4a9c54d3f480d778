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
namespace gla {

// passes 2 and 4: the two parities through the transform together (2, one
// LDS buffer each) or one after the other through one buffer (1: half the
// LDS per workgroup, so more workgroups per CU) -- per element type and
// length from ytune
template <typename R, int N>
constexpr int yk_of()
{
    return ytune((int)sizeof(cpx<R>), N);
}

// A partition of a volume as pass 1 reads it (source) or pass 5 writes it
// (target): cell (i1, i2, i3) of the partition is cell (i_d * d_d + o_d) of
// the full volume of f_d cells per axis (genPrt / mrgPrt,
// src/glaActSup.jl:4-18, 939-953).  Unpartitioned: f = the volume, d = 1,
// o = 0.
struct Part {
    int f[3], d[3], o[3];
};


// (i2, i3, c) of row `pen` of a workgroup whose first row is `row0`, rows
// ordered (i2 fastest, then i3, then c) over n2 x n3 x 3: the divisions are
// the workgroup's own (uniform, scalar), each lane then steps on by pen rows
__device__ __forceinline__ void row_coords(int row0, int pen, int n2, int n3, int& i2, int& i3, int& c)
{
    const unsigned r0 = (unsigned)row0, u2 = (unsigned)n2, u3 = (unsigned)n3;
    int a = (int)(r0 % u2), q = (int)(r0 / u2);
    int b = (int)((unsigned)q % u3), cc = (int)((unsigned)q / u3);
    a += pen;
    while (a >= n2) {
        a -= n2;
        if (++b == n3) {
            b = 0;
            ++cc;
        }
    }
    i2 = a;
    i3 = b;
    c = cc;
}

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

// The two parities of an axis-1 pencil go through the transform stages
// together (2: two LDS buffers; a stage with fewer butterflies than threads
// then deals both parities' butterflies over all lanes) or one after the
// other through one buffer (1: half the LDS per workgroup, so more
// workgroups fit a CU) -- per element type and length from xtune.
template <typename R, int N>
constexpr int xk_of()
{
    return xtune((int)sizeof(cpx<R>), N);
}

// --------------------------------------------------------------------------
// Pass 1: x-forward on contiguous rows.  One pencil = one (i2,i3,c) row of X.
// --------------------------------------------------------------------------
template <typename R, int N, int TP, int PEN, int MINW>
__global__ __launch_bounds__(TP * PEN, MINW) void k_fwd_x(const cpx<R>* __restrict__ X, cpx<R>* __restrict__ W1,
                                                          int S1, int S2, int S3, int Y1, int Z,
                                                          int64_t xs, int64_t ws,
                                                          const cpx<R>* __restrict__ phz,
                                                          const cpx<R>* __restrict__ tw, Part xp)
{
    constexpr int E = N / TP;
    // one padding slot every 8 positions (stride-8 stage writes hit distinct banks)
    constexpr int PADSH = 3;
    constexpr int PS = N + (N >> PADSH) + 1;
    // rows are whole pencils inside one wave when TP divides 64
    constexpr int XSYNC = (64 % TP == 0) ? 1 : 0;
    __shared__ cpx<R> lds[xk_of<R, N>() * PEN * PS];
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
    int i2, i3, c;
    row_coords(blockIdx.x * PEN, act ? pen : 0, S2, S3, i2, i3, c);
    // the row of source partition xp (the whole volume's row without
    // partitions: xp.f = S, xp.d = 1, xp.o = 0)
    const cpx<R>* xr = X + (int64_t)blockIdx.y * xs +
                       ((((int64_t)c * xp.f[2] + (int64_t)i3 * xp.d[2] + xp.o[2]) * xp.f[1] + (int64_t)i2 * xp.d[1] +
                         xp.o[1]) * xp.f[0] + xp.o[0]);
    const int xd = xp.d[0];
    cpx<R>* wr = W1 + (int64_t)blockIdx.y * ws + (((int64_t)c * Z + i3) * Y1 + i2) * (2 * N);
    cpx<R> v[2][E];   // even, odd parity
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
        if (act && n < S1) a = xr[(int64_t)n * xd];
        if (act && n + N < S1) b = xr[(int64_t)(n + N) * xd];
        v[0][e] = a + b;
        v[1][e] = phz[n] * (a - b);
    }
    if constexpr (xk_of<R, N>() == 2) {
        PencilFft<R, N, TP, 1, PS, false, XSYNC, 2, PEN * PS, PADSH>::run(v, lds, t, pen, tw);
    } else {
        PencilFft<R, N, TP, 1, PS, false, XSYNC, 1, 0, PADSH>::run(v[0], lds, t, pen, tw);
        lds_sync<XSYNC>();
        PencilFft<R, N, TP, 1, PS, false, XSYNC, 1, 0, PADSH>::run(v[1], lds, t, pen, tw);
    }
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            wr[t + TP * e] = v[0][e];
            wr[N + t + TP * e] = v[1][e];
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
                                                          const cpx<R>* __restrict__ chi, int64_t chis, Part yp)
{
    // axpby 0: y = G x;  1: y = alpha G x + beta y;  2: y = x - chi (G x),
    // chi one value per target cell (per right-hand side when chis > 0)
    constexpr int E = N / TP;
    constexpr int PADSH = 3;
    constexpr int PS = N + (N >> PADSH) + 1;
    constexpr int XSYNC = (64 % TP == 0) ? 1 : 0;
    __shared__ cpx<R> lds[xk_of<R, N>() * PEN * PS];
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
    int i2, i3, c;
    row_coords(blockIdx.x * PEN, act ? pen : 0, T2, T3, i2, i3, c);
    const cpx<R>* wr = W1 + (int64_t)blockIdx.y * ws + (((int64_t)c * Z + i3) * Y1 + i2) * (2 * N);
    // the row of target partition yp (the whole volume's row without
    // partitions); the Lippmann-Schwinger form (axpby 2) is for self
    // operators, never partitioned
    cpx<R>* yr = Y + (int64_t)blockIdx.y * ys +
                 ((((int64_t)c * yp.f[2] + (int64_t)i3 * yp.d[2] + yp.o[2]) * yp.f[1] + (int64_t)i2 * yp.d[1] + yp.o[1]) *
                      yp.f[0] + yp.o[0]);
    const int yd = yp.d[0];
    const int64_t cell0 = ((int64_t)i3 * T2 + i2) * T1;
    cpx<R> v[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = t + TP * e;
        v[0][e] = act ? wr[k] : cpx<R>{R(0), R(0)};
        v[1][e] = act ? wr[N + k] : cpx<R>{R(0), R(0)};
    }
    if constexpr (xk_of<R, N>() == 2) {
        PencilFft<R, N, TP, 1, PS, true, XSYNC, 2, PEN * PS, PADSH>::run(v, lds, t, pen, tw);
    } else {
        PencilFft<R, N, TP, 1, PS, true, XSYNC, 1, 0, PADSH>::run(v[0], lds, t, pen, tw);
        lds_sync<XSYNC>();
        PencilFft<R, N, TP, 1, PS, true, XSYNC, 1, 0, PADSH>::run(v[1], lds, t, pen, tw);
    }
    if (!act) return;
    const int top = T1 < N ? T1 : N;
    const bool rd = axpby == 1 && !(beta.x == R(0) && beta.y == R(0));
    const cpx<R>* xr = axpby == 2 ? lsx + (int64_t)blockIdx.y * lsxs + (int64_t)row * T1 : nullptr;
    const cpx<R>* cr = axpby == 2 ? chi + (int64_t)blockIdx.y * chis + cell0 : nullptr;
    auto epilogue = [&](cpx<R> r, int n) {
        const int64_t ny = (int64_t)n * yd;
        if (axpby == 1) r = alpha * r + (rd ? beta * yr[ny] : cpx<R>{R(0), R(0)});
        else if (axpby == 2) r = xr[n] - cr[n] * r;
        yr[ny] = r;
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
    __shared__ cpx<R> lds[yk_of<R, N>() * PX * N];
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
    // row k + t of this thread's column: workgroup-uniform base plus a 32-bit
    // per-lane byte offset (no 64-bit multiply per access)
    const uint32_t lob = ((uint32_t)xl + (uint32_t)t * (uint32_t)L1) * (uint32_t)sizeof(cpx<R>);
    const cpx<R>* const su = W1 + (int64_t)blockIdx.z * w1s + ((int64_t)c * Z + i3) * Y1 * L1 + (int64_t)blockIdx.x * PX;
    cpx<R>* const du = W2 + (int64_t)blockIdx.z * w2s + ((int64_t)c * Z + i3) * (int64_t)(2 * N) * L1 +
                       (int64_t)blockIdx.x * PX;
    auto sp = [&](int k) { return reinterpret_cast<const cpx<R>*>(reinterpret_cast<const char*>(su + (int64_t)k * L1) + lob); };
    auto dp = [&](int k) { return reinterpret_cast<cpx<R>*>(reinterpret_cast<char*>(du + (int64_t)k * L1) + lob); };
    cpx<R> v[2][E];
    // whole bundle inside L1 and S2 >= N: the first-half loads need no guards
    if ((int)blockIdx.x * PX + PX <= L1 && S2 >= N) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[0][e] = *sp(TP * e);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + TP * e;
            v[0][e] = (act && n < S2) ? *sp(TP * e) : cpx<R>{R(0), R(0)};
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TP * e;
        const cpx<R> a = v[0][e];
        cpx<R> b = {R(0), R(0)};
        if (act && n + N < S2) b = *sp(TP * e + N);
        v[0][e] = a + b;
        v[1][e] = phz[n] * (a - b);
    }
    if constexpr (yk_of<R, N>() == 2) {
        PencilFft<R, N, TP, PX, 1, false, 0, 2, PX * N>::run(v, lds, t, xl, tw);
    } else {
        PencilFft<R, N, TP, PX, 1, false>::run(v[0], lds, t, xl, tw);
        PencilFft<R, N, TP, PX, 1, false>::run(v[1], lds, t, xl, tw);
    }
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            *dp(TP * e) = v[0][e];
            *dp(N + TP * e) = v[1][e];
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
    __shared__ cpx<R> lds[yk_of<R, N>() * PX * N];
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
    // row k + t of this thread's column: workgroup-uniform base plus a 32-bit
    // per-lane byte offset (no 64-bit multiply per access)
    const uint32_t lob = ((uint32_t)xl + (uint32_t)t * (uint32_t)L1) * (uint32_t)sizeof(cpx<R>);
    const cpx<R>* const su = W2 + (int64_t)blockIdx.z * w2s + ((int64_t)c * Z + i3) * (int64_t)(2 * N) * L1 +
                             (int64_t)blockIdx.x * PX;
    cpx<R>* const du = W1 + (int64_t)blockIdx.z * w1s + ((int64_t)c * Z + i3) * Y1 * L1 + (int64_t)blockIdx.x * PX;
    auto sp = [&](int k) { return reinterpret_cast<const cpx<R>*>(reinterpret_cast<const char*>(su + (int64_t)k * L1) + lob); };
    auto dp = [&](int k) { return reinterpret_cast<cpx<R>*>(reinterpret_cast<char*>(du + (int64_t)k * L1) + lob); };
    cpx<R> v[2][E];
    // whole bundle inside L1: the loads need no guards
    if ((int)blockIdx.x * PX + PX <= L1) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            v[0][e] = *sp(TP * e);
            v[1][e] = *sp(N + TP * e);
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            v[0][e] = act ? *sp(TP * e) : cpx<R>{R(0), R(0)};
            v[1][e] = act ? *sp(N + TP * e) : cpx<R>{R(0), R(0)};
        }
    }
    if constexpr (yk_of<R, N>() == 2) {
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
        if (n < top) *dp(TP * e) = v[0][e] + po;
        if (n + N < T2) *dp(TP * e + N) = v[0][e] - po;
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
    const int64_t zs = (int64_t)L2 * L1;
    // Column addressing: a workgroup-uniform base (scalar registers) plus one
    // 32-bit per-lane byte offset (xl + t*zs elements; the launcher checks
    // that TP*L1*L2 elements fit in 32-bit byte offsets), so no access needs
    // a 64-bit multiply.  cp(c, k) is element (c, k + t) of this thread's
    // column m1, k uniform.
    cpx<R>* const wgc = W2 + (int64_t)rhs * w2s + (int64_t)m2 * L1 + (int64_t)(tile & 0xffff) * PX;
    const uint32_t lob = ((uint32_t)xl + (uint32_t)t * (uint32_t)zs) * (uint32_t)sizeof(cpx<R>);
    auto cp = [&](int c, int k) -> cpx<R>* {
        return reinterpret_cast<cpx<R>*>(reinterpret_cast<char*>(wgc + ((int64_t)c * cs + (int64_t)k * zs)) + lob);
    };

    // RELOAD: read the column again for the odd parity (an L2 hit) instead of
    // keeping it in registers through the even parity's transforms
    constexpr int XE = RELOAD ? 1 : E;
    // full: every lane of the workgroup reads rows 0..N-1 of an existing
    // column (whole tile inside L1, S3 >= N), so the first-half loads need no
    // guards and issue back to back; guarded loads otherwise
    const bool full = (tile & 0xffff) * PX + PX <= L1 && S3 >= N;
    auto load_first_half = [&](cpx<R>(&w)[3][E]) {
        if (full) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < E; ++e) w[c][e] = *cp(c, TP * e);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int n = t + TP * e;
                    w[c][e] = (act && n < S3) ? *cp(c, TP * e) : cpx<R>{R(0), R(0)};
                }
        }
    };
    // GG bit 0: the transform's root table and the split phases staged in
    // LDS once per workgroup (their reads inside the stages and the split /
    // merge then wait on LDS, not on the L2).  With the first G^ point
    // prefetched (GG bit 2) the table is staged before the column loads are
    // issued, so the wait for the table's loads does not also wait for the
    // column and the prefetch stays in flight through the forward transform
    // (128^3 ComplexF32 pass 3: 162 -> 158 us); otherwise after them, so the
    // column loads overlap the staging.
    constexpr bool TAB_FIRST = (GG & 4) != 0;
    __shared__ cpx<R> twl[(GG & 1) ? 2 * N : 1];
    auto stage_tab = [&]() {
        if constexpr ((GG & 1) != 0) {
            for (int i = threadIdx.x; i < N; i += TP * PX) {
                twl[i] = tw[i];
                twl[N + i] = phz[i];
            }
            __syncthreads();
        }
    };
    if constexpr (TAB_FIRST) stage_tab();
    cpx<R> xa[3][XE];
    cpx<R> xb[MODE == 1 ? 3 : 1][MODE == 1 ? XE : 1];
    if constexpr (RELOAD == 0) {
        load_first_half(xa);
        if constexpr (MODE == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int n = t + TP * e;
                    xb[c][e] = (act && n + N < S3) ? *cp(c, TP * e + N) : cpx<R>{R(0), R(0)};
                }
        }
    }

    // per-column branch parities and stored indices on axes 1 and 2
    const int p1 = m1 >= a1.B ? 1 : 0, k1 = m1 - p1 * a1.B;
    const int p2 = m2 >= a2.B ? 1 : 0, k2 = m2 - p2 * a2.B;
    int j1, s1, j2, s2;
    axis_lookup(a1, p1, act ? k1 : 0, j1, s1);
    axis_lookup(a2, p2, k2, j2, s2);

    if constexpr (!TAB_FIRST) stage_tab();
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
    // issued by every lane (an inactive lane's column index is clamped to a
    // stored entry), so the wait for the column data before the forward
    // transform can leave these loads in flight
    auto pre_g = [&](int p3) {
        if constexpr (GPRE) g_fetch<R, TP>(gbase(p3), uu, a3, p3, t, 0, g0, g0s);
    };
    auto mul_g = [&](cpx<R>(&w)[3][E], int p3) {
        g_multiply<R, E, TP, GPIPE>(w, gbase(p3), uu, a3, p3, t, s1, s2, act, GPRE ? &g0 : nullptr, g0s);
    };
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
        load_first_half(wv[0]);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int n = t + TP * e;
                const cpx<R> a = wv[0][c][e];
                if constexpr (MODE == 1) {
                    const cpx<R> b = (act && n + N < S3) ? *cp(c, TP * e + N) : cpx<R>{R(0), R(0)};
                    wv[0][c][e] = a + b;
                    wv[1][c][e] = phu[n] * (a - b);
                } else {
                    wv[1][c][e] = phu[n] * a;   // no spill-over half
                }
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
                if (n < top) *cp(c, TP * e) = wv[0][c][e] + po;
                if constexpr (MODE == 2)
                    if (n + N < T3) *cp(c, TP * e + N) = wv[0][c][e] - po;
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
                    a = (act && n < S3) ? *cp(c, TP * e) : cpx<R>{R(0), R(0)};
                    if constexpr (MODE == 1)
                        b = (act && n + N < S3) ? *cp(c, TP * e + N) : cpx<R>{R(0), R(0)};
                } else {
                    a = xa[c][e];
                    if constexpr (MODE == 1) b = xb[c][e];
                }
                if constexpr (MODE == 1) w[c][e] = p3 == 0 ? a + b : phu[n] * (a - b);
                else w[c][e] = p3 == 0 ? a : phu[n] * a;   // no spill-over half
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
            if (n < top) *cp(c, TP * e) = acc[c][e];
            if constexpr (MODE == 2)
                if (n + N < T3) *cp(c, TP * e + N) = accb[c][e];
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
