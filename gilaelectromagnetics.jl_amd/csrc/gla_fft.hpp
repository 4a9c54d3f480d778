// Device FFT engine for the Green-function matvec (gfx950).
//
// A "pencil" is one 1-D transform of length N.  TP threads cooperate on a
// pencil and thread t holds the elements n = t + TP*e, e < E = N/TP, in
// registers ("natural residue layout").  The transform is a mixed-radix
// Stockham sequence; stage s with radix r and running span ns reads
// x[j + q*N/r], multiplies by w^{q*(j mod ns)} (w = e^{-2 pi i/(ns r)} forward),
// applies DFT_r and writes y[(j/ns)*ns*r + (j mod ns) + q*ns].  When the first
// radix divides E the first stage reads straight from registers, when the last
// radix divides E the last stage writes straight back to registers in the same
// residue layout, so a two-stage transform costs one LDS exchange.  Middle
// stages go through LDS.  All loops are compile-time unrolled so the register
// arrays never spill to scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gla_tune.hpp"

namespace gla {

template <typename R>
struct cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> mk(R a, R b) { return cpx<R>{a, b}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator*(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> operator*(R s, cpx<R> a) { return {s * a.x, s * a.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmulc(cpx<R> a, cpx<R> b)   // a * conj(b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> conjg(cpx<R> a) { return {a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV>
__device__ __forceinline__ cpx<R> mul_mi(cpx<R> a)
{
    if (INV) return {-a.y, a.x};
    return {a.y, -a.x};
}

// ---------------------------------------------------------------------------
// compile-time radix plan
// ---------------------------------------------------------------------------
struct Plan {
    int n;
    int nst;
    int rad[24];
};

__host__ __device__ constexpr int smallest_factor(int m)
{
    for (int f = 2; f * f <= m; ++f)
        if (m % f == 0) return f;
    return m;
}

// Factor N into radices from {16,8,4,2,3,5,other primes}; the first radix is
// the largest supported radix dividing both N and E, the last the largest
// dividing what remains and E, the middle ones the largest remaining factors
// (pol 0: fewest LDS exchanges) or, where one exists, the largest that also
// divides E (pol 1: a middle stage of radix r runs E / r butterflies per
// thread, so every thread stays busy, at the cost of more exchanges).
__host__ __device__ constexpr Plan make_plan(int N, int E, int pol = 0)
{
    Plan p{N, 0, {}};
    int m = N;
    int mid[24] = {};
    int nmid = 0;
    int first = 0, last = 0;
    constexpr int pref[6] = {16, 8, 4, 5, 3, 2};
    for (int i = 0; i < 6 && !first; ++i)
        if (m % pref[i] == 0 && E % pref[i] == 0) first = pref[i];
    if (first) m /= first;
    for (int i = 0; i < 6 && !last && m > 1; ++i)
        if (m % pref[i] == 0 && E % pref[i] == 0) last = pref[i];
    if (last) m /= last;
    while (m > 1) {
        int r = 0;
        if (pol == 1)
            for (int i = 0; i < 6 && !r; ++i)
                if (m % pref[i] == 0 && E % pref[i] == 0) r = pref[i];
        for (int i = 0; i < 6 && !r; ++i)
            if (m % pref[i] == 0) r = pref[i];
        if (!r) r = smallest_factor(m);
        mid[nmid++] = r;
        m /= r;
    }
    if (first) p.rad[p.nst++] = first;
    for (int i = 0; i < nmid; ++i) p.rad[p.nst++] = mid[i];
    if (last) p.rad[p.nst++] = last;
    return p;
}

// ---------------------------------------------------------------------------
// radix codelets (in place on a[0..r-1]); w = e^{-+2 pi i / r}
// ---------------------------------------------------------------------------
template <typename R, bool INV>
__device__ __forceinline__ void dft2(cpx<R>* a)
{
    cpx<R> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
}

template <typename R, bool INV>
__device__ __forceinline__ void dft3(cpx<R>* a)
{
    const R c1 = R(-0.5);
    const R s1 = INV ? R(0.86602540378443864676) : R(-0.86602540378443864676);
    cpx<R> s = a[1] + a[2], d = a[1] - a[2];
    cpx<R> u = a[0] + c1 * s;
    cpx<R> v = {-s1 * d.y, s1 * d.x};
    a[0] = a[0] + s;
    a[1] = u + v;
    a[2] = u - v;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4(cpx<R>* a)
{
    cpx<R> s02 = a[0] + a[2], d02 = a[0] - a[2];
    cpx<R> s13 = a[1] + a[3], d13 = mul_mi<R, INV>(a[1] - a[3]);
    a[0] = s02 + s13;
    a[1] = d02 + d13;
    a[2] = s02 - s13;
    a[3] = d02 - d13;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft5(cpx<R>* a)
{
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = INV ? R(0.95105651629515357212) : R(-0.95105651629515357212);
    const R s2 = INV ? R(0.58778525229247312917) : R(-0.58778525229247312917);
    cpx<R> p1 = a[1] + a[4], m1 = a[1] - a[4];
    cpx<R> p2 = a[2] + a[3], m2 = a[2] - a[3];
    cpx<R> u1 = a[0] + c1 * p1 + c2 * p2;
    cpx<R> u2 = a[0] + c2 * p1 + c1 * p2;
    cpx<R> v1 = s1 * m1 + s2 * m2;          // times i
    cpx<R> v2 = s2 * m1 - s1 * m2;          // times i
    a[0] = a[0] + p1 + p2;
    a[1] = u1 + cpx<R>{-v1.y, v1.x};
    a[4] = u1 - cpx<R>{-v1.y, v1.x};
    a[2] = u2 + cpx<R>{-v2.y, v2.x};
    a[3] = u2 - cpx<R>{-v2.y, v2.x};
}

template <typename R, bool INV>
__device__ __forceinline__ void dft8(cpx<R>* a)
{
    const R h = R(0.70710678118654752440);
    cpx<R> b[4] = {a[0], a[2], a[4], a[6]};
    cpx<R> c[4] = {a[1], a[3], a[5], a[7]};
    dft4<R, INV>(b);
    dft4<R, INV>(c);
    // twiddles w8^k, k = 1..3
    cpx<R> c1 = c[1], c2 = c[2], c3 = c[3];
    if (!INV) {
        c[1] = {h * (c1.x + c1.y), h * (c1.y - c1.x)};
        c[2] = {c2.y, -c2.x};
        c[3] = {h * (c3.y - c3.x), -h * (c3.x + c3.y)};
    } else {
        c[1] = {h * (c1.x - c1.y), h * (c1.x + c1.y)};
        c[2] = {-c2.y, c2.x};
        c[3] = {-h * (c3.x + c3.y), h * (c3.x - c3.y)};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = b[k] + c[k];
        a[k + 4] = b[k] - c[k];
    }
}

template <typename R, bool INV>
__device__ __forceinline__ void dft16(cpx<R>* a)
{
    // 16 = 4 x 4: inner DFT4 over stride-4 elements, twiddle w16^{q1*k2}, outer DFT4
    const R c8 = R(0.92387953251128675613), s8 = R(0.38268343236508977173), h = R(0.70710678118654752440);
    cpx<R> b[4][4];
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) b[q1][q2] = a[q1 + 4 * q2];
        dft4<R, INV>(b[q1]);
    }
    // twiddle b[q1][k2] *= w16^{q1 k2}
    const R sg = INV ? R(1) : R(-1);
    const cpx<R> w[10] = {
        {R(1), R(0)}, {c8, sg * s8}, {h, sg * h}, {s8, sg * c8}, {R(0), sg},
        {-s8, sg * c8}, {-h, sg * h}, {-c8, sg * s8}, {R(-1), R(0)}, {-c8, -sg * s8}};
#pragma unroll
    for (int q1 = 1; q1 < 4; ++q1)
#pragma unroll
        for (int k2 = 1; k2 < 4; ++k2) b[q1][k2] = b[q1][k2] * w[q1 * k2];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        cpx<R> c[4] = {b[0][k2], b[1][k2], b[2][k2], b[3][k2]};
        dft4<R, INV>(c);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[k2 + 4 * k1] = c[k1];
    }
}

// generic odd radix: direct sum over the pairs (e, RAD - e), whose roots are
// conjugate: with s_e = a_e + a_{RAD-e}, d_e = a_e - a_{RAD-e} and
// w^{eq} = C + i S,
//   out[q] = a_0 + sum_e (s_e C + i d_e S),  out[RAD-q] = a_0 + sum_e (s_e C - i d_e S),
// so each (q, e) costs two real-by-complex products for two outputs.  The
// (RAD-1)/2 distinct roots come from the length-N table once per call.
template <typename R, bool INV, int RAD, int N>
__device__ __forceinline__ void dftg(cpx<R>* a, const cpx<R>* __restrict__ tw)
{
    constexpr int H = (RAD - 1) / 2;
    static_assert(RAD % 2 == 1, "the generic radix is odd");
    cpx<R> rt[H + 1];
#pragma unroll
    for (int k = 1; k <= H; ++k) rt[k] = tw[k * (N / RAD)];
    cpx<R> sm[H + 1], df[H + 1];
    cpx<R> o0 = a[0];
#pragma unroll
    for (int e = 1; e <= H; ++e) {
        sm[e] = a[e] + a[RAD - e];
        df[e] = a[e] - a[RAD - e];
        o0 = o0 + sm[e];
    }
    cpx<R> o[RAD];
    o[0] = o0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        cpx<R> X = a[0], Y = {R(0), R(0)};
#pragma unroll
        for (int e = 1; e <= H; ++e) {
            const int k = (e * q) % RAD;
            // forward root w^k = e^{-2 pi i k / RAD}; w^k = conj(w^{RAD-k})
            const R C = k <= H ? rt[k].x : rt[RAD - k].x;
            const R S = (k <= H ? rt[k].y : -rt[RAD - k].y) * (INV ? R(-1) : R(1));
            X = X + C * sm[e];
            Y = Y + S * df[e];
        }
        const cpx<R> iY = {-Y.y, Y.x};
        o[q] = X + iY;
        o[RAD - q] = X - iY;
    }
#pragma unroll
    for (int q = 0; q < RAD; ++q) a[q] = o[q];
}

template <typename R, bool INV, int RAD, int N>
__device__ __forceinline__ void dft(cpx<R>* a, const cpx<R>* __restrict__ tw)
{
    if constexpr (RAD == 2) dft2<R, INV>(a);
    else if constexpr (RAD == 3) dft3<R, INV>(a);
    else if constexpr (RAD == 4) dft4<R, INV>(a);
    else if constexpr (RAD == 5) dft5<R, INV>(a);
    else if constexpr (RAD == 8) dft8<R, INV>(a);
    else if constexpr (RAD == 16) dft16<R, INV>(a);
    else dftg<R, INV, RAD, N>(a, tw);
}

// ---------------------------------------------------------------------------
// Twiddled butterflies with the twiddle products fused into the butterfly's
// additions: for a + w x and a - w x, s = a + w x takes two FMAs per
// component and d = 2a - s one, against a complex product and an add and a
// subtract (eight operations) separately.  a[0] is untwiddled, a[q] is
// multiplied by w[q] (conjugated for the inverse).  Per element type:
// GLA_FUSED_TW_F (ComplexF32) and GLA_FUSED_TW_D (ComplexF64), 0 keeps the
// separate products.  Measured at 128^3: ComplexF64 pass 3 317 -> 313 us
// with them; ComplexF32 pass 3 159 -> 169 us (fewer instructions, but
// longer dependent chains at the same occupancy), so off there.
// ---------------------------------------------------------------------------
// stages with fewer butterflies than threads deal the (pencil, butterfly)
// tasks of K pencils over all threads (PencilFft::stage_lds_spread); 2 also
// deals stages whose butterflies do not divide evenly over the threads when
// that lowers the most tasks one thread runs; 0 keeps threads t < NB doing
// every pencil of butterfly t.  Per element
// type: mode 2 measured faster in ComplexF64 (pass 3 at 96^3 165 -> 160 us,
// 192^3 1326 -> 1288 us) and not in ComplexF32 (the external 256x34x34 in
// 256x256x128 pair's pass 3 363 -> 379 us).
#ifndef GLA_SPREAD_F
#define GLA_SPREAD_F 1
#endif
#ifndef GLA_SPREAD_D
#define GLA_SPREAD_D 2
#endif
#ifndef GLA_FUSED_TW_F
#define GLA_FUSED_TW_F 0
#endif
#ifndef GLA_FUSED_TW_D
#define GLA_FUSED_TW_D 1
#endif

template <typename R, bool INV>
__device__ __forceinline__ void fma_pair(cpx<R> a, cpx<R> x, cpx<R> w, cpx<R>& s, cpx<R>& d)
{
    const R wy = INV ? -w.y : w.y;
    s.x = fma(x.x, w.x, fma(-x.y, wy, a.x));
    s.y = fma(x.x, wy, fma(x.y, w.x, a.y));
    d.x = fma(R(2), a.x, -s.x);
    d.y = fma(R(2), a.y, -s.y);
}

template <typename R, bool INV>
__device__ __forceinline__ cpx<R> twid_(cpx<R> a, cpx<R> w)
{
    return INV ? cmulc(a, w) : a * w;
}

// dft4 of (a0, w1 a1, w2 a2, w3 a3)
template <typename R, bool INV>
__device__ __forceinline__ void dft4_tw(cpx<R>* a, cpx<R> w1, cpx<R> w2, cpx<R> w3)
{
    cpx<R> s02, d02, s13, d13;
    fma_pair<R, INV>(a[0], a[2], w2, s02, d02);
    fma_pair<R, INV>(twid_<R, INV>(a[1], w1), a[3], w3, s13, d13);
    d13 = mul_mi<R, INV>(d13);
    a[0] = s02 + s13;
    a[1] = d02 + d13;
    a[2] = s02 - s13;
    a[3] = d02 - d13;
}

// dft8 of (a0, w1 a1, ..., w7 a7): the two inner dft4 take their twiddles
// fused, then the w8 rotations and the outer butterflies of dft8
template <typename R, bool INV>
__device__ __forceinline__ void dft8_tw(cpx<R>* a, const cpx<R>* w)
{
    const R h = R(0.70710678118654752440);
    cpx<R> b[4] = {a[0], a[2], a[4], a[6]};
    cpx<R> c[4] = {twid_<R, INV>(a[1], w[1]), a[3], a[5], a[7]};
    dft4_tw<R, INV>(b, w[2], w[4], w[6]);
    // c = dft4 of (w1 a1, w3 a3, w5 a5, w7 a7): its first element is already
    // twiddled, the others fused
    {
        cpx<R> s02, d02, s13, d13;
        fma_pair<R, INV>(c[0], c[2], w[5], s02, d02);
        fma_pair<R, INV>(twid_<R, INV>(c[1], w[3]), c[3], w[7], s13, d13);
        d13 = mul_mi<R, INV>(d13);
        c[0] = s02 + s13;
        c[1] = d02 + d13;
        c[2] = s02 - s13;
        c[3] = d02 - d13;
    }
    cpx<R> c1 = c[1], c2 = c[2], c3 = c[3];
    if (!INV) {
        c[1] = {h * (c1.x + c1.y), h * (c1.y - c1.x)};
        c[2] = {c2.y, -c2.x};
        c[3] = {h * (c3.y - c3.x), -h * (c3.x + c3.y)};
    } else {
        c[1] = {h * (c1.x - c1.y), h * (c1.x + c1.y)};
        c[2] = {-c2.y, c2.x};
        c[3] = {-h * (c3.x + c3.y), h * (c3.x - c3.y)};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = b[k] + c[k];
        a[k + 4] = b[k] - c[k];
    }
}

// radices with a fused twiddled butterfly for element type R
template <typename R, int RAD>
constexpr bool fused_tw()
{
    return (sizeof(R) == 8 ? GLA_FUSED_TW_D : GLA_FUSED_TW_F) && (RAD == 2 || RAD == 4 || RAD == 8);
}

// a[q] *= w[q] (q >= 1), then the radix-RAD butterfly (the radices of
// fused_tw; the others twiddle, then call dft)
template <typename R, bool INV, int RAD>
__device__ __forceinline__ void dft_tw(cpx<R>* a, const cpx<R>* w)
{
    static_assert(RAD == 2 || RAD == 4 || RAD == 8, "fused twiddled butterflies: radix 2, 4, 8");
    if constexpr (RAD == 2) {
        cpx<R> s, d;
        fma_pair<R, INV>(a[0], a[1], w[1], s, d);
        a[0] = s;
        a[1] = d;
    } else if constexpr (RAD == 4) {
        dft4_tw<R, INV>(a, w[1], w[2], w[3]);
    } else {
        dft8_tw<R, INV>(a, w);
    }
}

// ---------------------------------------------------------------------------
// LDS exchange fences.  SYNC 0: workgroup barrier (a pencil's threads span
// several waves).  SYNC 1: the pencil's threads are lanes of one wave, whose
// LDS operations complete in issue order, so a compiler fence ordering the
// writes before the reads is enough.
// ---------------------------------------------------------------------------
template <int SYNC>
__device__ __forceinline__ void lds_sync()
{
    if constexpr (SYNC == 0) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// The pencil transform.
//   v    : K independent pencils per thread, E = N/TP register elements each,
//          v[k][e] = x_k[t + TP*e]
//   lds  : WG scratch; element pos of pencil k lives at
//          lds[k*KS + pos*NS + pen*PS]
//   tw   : table tw[n] = e^{-2 pi i n / N}, n < N (forward sign)
// The K pencils go through each stage together: one twiddle load serves all
// of them and one pair of fences per stage covers all of them.  All threads
// of the workgroup (SYNC 0) or of the wave (SYNC 1) must call with the same
// template arguments.  A masked pencil beyond the data still takes part.
// ---------------------------------------------------------------------------
template <typename R, int N, int TP, int NS, int PS, bool INV, int SYNC = 0, int K = 1, int KS = 0, int PADSH = 0>
struct PencilFft {
    // LDS slot of position pos: with PADSH > 0 one slot is skipped every
    // 2^PADSH positions, so the stride-RAD writes of a stage spread over the
    // banks (used when positions of one pencil are contiguous, NS == 1)
    static __device__ __forceinline__ int slot(int pos)
    {
        if constexpr (PADSH > 0) return pos + (pos >> PADSH);
        return pos;
    }
    static constexpr int E = N / TP;
    static_assert(N % TP == 0, "TP must divide N");
    static constexpr Plan P = make_plan(N, E, ptune((int)sizeof(cpx<R>), N));
    static constexpr bool FIRST_REG = (P.nst > 0) && (E % P.rad[0] == 0);
    static constexpr bool LAST_REG = (P.nst > 1) && (E % P.rad[P.nst - 1] == 0);

    static __device__ __forceinline__ cpx<R> twid(cpx<R> a, cpx<R> w) { return INV ? cmulc(a, w) : a * w; }

    // one stage reading LDS and writing LDS (any thread mapping)
    template <int S>
    static __device__ __forceinline__ void stage_lds(cpx<R>* lds, int t, int pen, const cpx<R>* tw)
    {
        constexpr int RAD = P.rad[S];
        constexpr int NB = N / RAD;                      // butterflies
        constexpr int NU = (NB + TP - 1) / TP;          // per thread (upper bound)
        constexpr int NSP = [] { int ns = 1; for (int i = 0; i < S; ++i) ns *= P.rad[i]; return ns; }();
        constexpr int SPR = sizeof(R) == 8 ? GLA_SPREAD_D : GLA_SPREAD_F;
        // mode 2: whenever dealing lowers the most tasks one thread runs
        constexpr bool LOWER = (K * NB + TP - 1) / TP < K * ((NB + TP - 1) / TP);
        if constexpr (SPR != 0 && K > 1 && (NB < TP || (SPR == 2 && LOWER))) {
            stage_lds_spread<RAD, NB, NSP>(lds, t, pen, tw);
            return;
        }
        cpx<R> a[K][NU][RAD];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int j = t + TP * u;
            if (NB % TP == 0 || j < NB) {
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int q = 0; q < RAD; ++q) a[k][u][q] = lds[k * KS + slot(j + q * NB) * NS + pen * PS];
            }
        }
        lds_sync<SYNC>();
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int j = t + TP * u;
            if (NB % TP == 0 || j < NB) {
                const int kk = j % NSP;
                const int base = (j / NSP) * NSP * RAD + kk;
                // twiddles, then per component the butterfly and its stores
                // (each component's stores issue before the next butterfly)
                cpx<R> w[RAD];
                if constexpr (NSP > 1) {
#pragma unroll
                    for (int q = 1; q < RAD; ++q) {
                        w[q] = tw[q * kk * (N / (NSP * RAD))];
                        if constexpr (!fused_tw<R, RAD>()) {
#pragma unroll
                            for (int k = 0; k < K; ++k) a[k][u][q] = twid(a[k][u][q], w[q]);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if constexpr (NSP > 1 && fused_tw<R, RAD>()) dft_tw<R, INV, RAD>(a[k][u], w);
                    else dft<R, INV, RAD, N>(a[k][u], tw);
#pragma unroll
                    for (int q = 0; q < RAD; ++q) lds[k * KS + slot(base + q * NSP) * NS + pen * PS] = a[k][u][q];
                }
            }
        }
        lds_sync<SYNC>();
    }

    // A stage with fewer butterflies than threads per pencil (NB < TP), or
    // (spread mode 2) butterflies that do not divide evenly over the threads,
    // and K pencils: the K * NB (pencil, butterfly) tasks are dealt over all
    // TP threads, task t + TP i to thread t, instead of thread t taking all K
    // pencils of butterflies t, t + TP, ...  Inside one wave this fills lanes
    // that would idle; across waves it evens out their work; each task reads
    // its own twiddles.
    template <int RAD, int NB, int NSP>
    static __device__ __forceinline__ void stage_lds_spread(cpx<R>* lds, int t, int pen, const cpx<R>* tw)
    {
        constexpr int NT = K * NB;
        constexpr int NUT = (NT + TP - 1) / TP;
        cpx<R> a[NUT][RAD];
#pragma unroll
        for (int i = 0; i < NUT; ++i) {
            const int tau = t + TP * i;
            if (NT % TP == 0 || tau < NT) {
                const int k = tau / NB, j = tau % NB;
#pragma unroll
                for (int q = 0; q < RAD; ++q) a[i][q] = lds[k * KS + slot(j + q * NB) * NS + pen * PS];
            }
        }
        lds_sync<SYNC>();
#pragma unroll
        for (int i = 0; i < NUT; ++i) {
            const int tau = t + TP * i;
            if (NT % TP == 0 || tau < NT) {
                const int k = tau / NB, j = tau % NB;
                const int kk = j % NSP;
                const int base = (j / NSP) * NSP * RAD + kk;
                cpx<R> w[RAD];
                if constexpr (NSP > 1) {
#pragma unroll
                    for (int q = 1; q < RAD; ++q) {
                        w[q] = tw[q * kk * (N / (NSP * RAD))];
                        if constexpr (!fused_tw<R, RAD>()) a[i][q] = twid(a[i][q], w[q]);
                    }
                }
                if constexpr (NSP > 1 && fused_tw<R, RAD>()) dft_tw<R, INV, RAD>(a[i], w);
                else dft<R, INV, RAD, N>(a[i], tw);
#pragma unroll
                for (int q = 0; q < RAD; ++q) lds[k * KS + slot(base + q * NSP) * NS + pen * PS] = a[i][q];
            }
        }
        lds_sync<SYNC>();
    }

    template <int S>
    static __device__ __forceinline__ void mid_stages(cpx<R>* lds, int t, int pen, const cpx<R>* tw)
    {
        constexpr int LASTS = LAST_REG ? P.nst - 1 : P.nst;
        if constexpr (S < LASTS) {
            stage_lds<S>(lds, t, pen, tw);
            mid_stages<S + 1>(lds, t, pen, tw);
        }
    }

    static __device__ __forceinline__ void run(cpx<R> (&v)[K][E], cpx<R>* lds, int t, int pen,
                                               const cpx<R>* __restrict__ tw)
    {
        if constexpr (P.nst == 0) {
            return;
        } else if constexpr (P.nst == 1 && FIRST_REG) {
            // whole transform inside one thread (E == N)
#pragma unroll
            for (int k = 0; k < K; ++k) dft<R, INV, P.rad[0], N>(v[k], tw);
        } else {
            // ---- first stage
            if constexpr (FIRST_REG) {
                constexpr int RAD = P.rad[0];
                constexpr int EU = E / RAD;   // butterflies per thread
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int u = 0; u < EU; ++u) {
                        cpx<R> a[RAD];
#pragma unroll
                        for (int q = 0; q < RAD; ++q) a[q] = v[k][u + q * EU];
                        dft<R, INV, RAD, N>(a, tw);
                        const int j = t + TP * u;    // butterfly index, span 1
#pragma unroll
                        for (int q = 0; q < RAD; ++q) lds[k * KS + slot(j * RAD + q) * NS + pen * PS] = a[q];
                    }
                lds_sync<SYNC>();
                mid_stages<1>(lds, t, pen, tw);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int e = 0; e < E; ++e) lds[k * KS + slot(t + TP * e) * NS + pen * PS] = v[k][e];
                lds_sync<SYNC>();
                mid_stages<0>(lds, t, pen, tw);
            }
            // ---- last stage
            if constexpr (LAST_REG) {
                constexpr int S = P.nst - 1;
                constexpr int RAD = P.rad[S];
                constexpr int EU = E / RAD;
                constexpr int NSP = N / RAD;
#pragma unroll
                for (int u = 0; u < EU; ++u) {
                    const int j = t + TP * u;
                    cpx<R> a[K][RAD];
#pragma unroll
                    for (int k = 0; k < K; ++k)
#pragma unroll
                        for (int q = 0; q < RAD; ++q) a[k][q] = lds[k * KS + slot(j + q * NSP) * NS + pen * PS];
                    cpx<R> w[RAD];
#pragma unroll
                    for (int q = 1; q < RAD; ++q) {
                        w[q] = tw[q * j];
                        if constexpr (!fused_tw<R, RAD>()) {
#pragma unroll
                            for (int k = 0; k < K; ++k) a[k][q] = twid(a[k][q], w[q]);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if constexpr (fused_tw<R, RAD>()) dft_tw<R, INV, RAD>(a[k], w);
                        else dft<R, INV, RAD, N>(a[k], tw);
#pragma unroll
                        for (int q = 0; q < RAD; ++q) v[k][u + q * EU] = a[k][q];
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
#pragma unroll
                    for (int e = 0; e < E; ++e) v[k][e] = lds[k * KS + slot(t + TP * e) * NS + pen * PS];
            }
            lds_sync<SYNC>();   // LDS free for the next transform
        }
    }

    // single pencil
    static __device__ __forceinline__ void run(cpx<R> (&v)[E], cpx<R>* lds, int t, int pen,
                                               const cpx<R>* __restrict__ tw)
    {
        static_assert(K == 1, "single-pencil form needs K == 1");
        run(reinterpret_cast<cpx<R>(&)[1][E]>(v), lds, t, pen, tw);
    }
};

}  // namespace gla
