// Operator construction from geometry on the GPU, in FP64:
//   1. the Green tensor per cell separation by face-pair surface integrals
//      (the reference's formulation: egoFunInn! / egoSrfAdp! / srfKer /
//      srfSum!, src/glaGen.jl:322-338, 531-598; cube faces and scalings of
//      cubFac / srfScl, src/glaInt.jl:284-328; scalar kernel sclEgo,
//      src/glaInt.jl:41-44).  Each face-pair integral is reduced to difference
//      coordinates (2-D for parallel faces, 3-D for perpendicular ones, with
//      the faces' overlap lengths as weights) and integrated with tensor
//      Gauss-Legendre rules on the pieces between the weights' kinks, to about
//      1e-14 relative -- in place of the reference's adaptive hcubature at
//      1e-6; for cells that touch on the cell grid the face pairs that
//      egoFunSng! marks (src/glaGen.jl:350-525) take the DIRECTFN singular
//      values of gla_singular.hip, as the reference's do;
//   2. the circulant embedding (egoToeCrc!, src/glaGen.jl:143-163, for self
//      operators; the separation grid of egoFunExt! / grdSel,
//      src/glaGen.jl:169-189, 632-643, for external ones);
//   3. the L-grid transform and even/odd branch extraction with symmetric
//      truncation (src/glaMemSup.jl:96-156), through the forward pass kernels.
// Only the upper-triangle components [xx,yy,zz,xy,xz,yz] are formed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "gla_launch.hpp"
#include "gla_setup.hpp"

namespace gla {
namespace {

constexpr int QMAX = 24;   // largest Gauss-Legendre order in the table

struct GreenCfg {
    double scl[3];          // cell size per axis (wavelengths); target cells when they differ
    double sclS[3];         // source cell size (equal to scl except for partitioned operators)
    double zr, zi;          // frqPhz
    double kr, ki;          // 2 pi frqPhz
    double cr, ci;          // 1 / (4 pi frqPhz^2)
    int qphase;             // smallest regular order for the cell's phase span
};

struct Cplx {
    double r, i;
};

// r g(r) = exp(2 pi i zeta r) / (4 pi zeta^2), smooth at r = 0
__device__ __forceinline__ Cplx rg(const GreenCfg& G, double r)
{
    double s, c;
    sincos(G.kr * r, &s, &c);
    const double m = G.ki == 0.0 ? 1.0 : exp(-G.ki * r);
    return Cplx{m * (c * G.cr - s * G.ci), m * (c * G.ci + s * G.cr)};
}

__device__ __forceinline__ const double* gl_tab(const double* gl, int q) { return gl + (q * (q - 1) / 2) * 2; }

// overlap length at difference w of two intervals of length h whose centres
// differ by c: the tent h - |w - c| on [c - h, c + h]
__device__ __forceinline__ double tent(double w, double c, double h) { return fmax(0.0, h - fabs(w - c)); }

constexpr double SNAP = 1e-12;   // relative: a breakpoint this close to zero is zero

// sorted, distinct breakpoints of [lo, hi]: the ends, the interior kink
// `mid` (NaN: none) and, when `zero` and it lies inside, the origin; values
// within SNAP * h of zero are zero.  Returns their count (at most 4).
__device__ int breaks(double lo, double mid, double hi, bool zero, double h, double* b)
{
    double c[4];
    int n = 0;
    c[n++] = lo;
    if (!isnan(mid)) c[n++] = mid;
    c[n++] = hi;
    if (zero && lo < 0.0 && hi > 0.0) c[n++] = 0.0;
    for (int k = 0; k < n; ++k)
        if (fabs(c[k]) < SNAP * h) c[k] = 0.0;
    for (int k = 1; k < n; ++k)
        for (int j = k; j > 0 && c[j] < c[j - 1]; --j) {
            const double t = c[j];
            c[j] = c[j - 1];
            c[j - 1] = t;
        }
    int m = 0;
    for (int k = 0; k < n; ++k)
        if (m == 0 || c[k] > b[m - 1]) b[m++] = c[k];
    return m;
}

// Face pair of parallel faces, in difference coordinates: the integral over
// (wa, wb) of g(sqrt(wa^2 + wb^2 + d^2)) tent_a(wa) tent_b(wb), tents centred
// at ca, cb with half widths ha, hb.  Pieces between the tents' kinks use a
// q x q rule; with d = 0 and the origin in the support, the pieces are also
// split at the origin and those with the origin at a corner are integrated in
// polar coordinates about it (r g(r) is smooth there).
__device__ Cplx par_pair(const GreenCfg& G, double d, double ca, double ha, double cb, double hb, int q, int qs,
                         const double* gl)
{
    const double hmax = fmax(ha, hb);
    if (fabs(d) < SNAP * hmax) d = 0.0;
    const bool sing = d == 0.0 && ca - ha <= SNAP * ha && ca + ha >= -SNAP * ha && cb - hb <= SNAP * hb &&
                      cb + hb >= -SNAP * hb;
    double ba[5], bb[5];
    const int na = breaks(ca - ha, ca, ca + ha, sing, ha, ba), nb = breaks(cb - hb, cb, cb + hb, sing, hb, bb);
    const double* x = gl_tab(gl, q);
    const double* xs = gl_tab(gl, qs);
    double ar = 0, ai = 0;
    for (int i = 0; i + 1 < na; ++i)
        for (int j = 0; j + 1 < nb; ++j) {
            const double a0 = ba[i], a1 = ba[i + 1], b0 = bb[j], b1 = bb[j + 1];
            if (sing && (a0 == 0.0 || a1 == 0.0) && (b0 == 0.0 || b1 == 0.0)) {
                const double su = a0 == 0.0 ? 1.0 : -1.0, sv = b0 == 0.0 ? 1.0 : -1.0;
                const double A = a1 - a0, B = b1 - b0, t0 = atan2(B, A);
                for (int tri = 0; tri < 2; ++tri) {
                    const double ta = tri ? t0 : 0.0, tb = tri ? 0.5 * M_PI : t0;
                    for (int k = 0; k < qs; ++k) {
                        const double t = ta + (tb - ta) * xs[2 * k], wt = (tb - ta) * xs[2 * k + 1];
                        double st, ct;
                        sincos(t, &st, &ct);
                        const double rm = tri ? B / st : A / ct;
                        for (int l = 0; l < qs; ++l) {
                            const double r = rm * xs[2 * l], w = wt * rm * xs[2 * l + 1];
                            const Cplx g = rg(G, r);
                            const double f = w * tent(su * r * ct, ca, ha) * tent(sv * r * st, cb, hb);
                            ar += f * g.r;
                            ai += f * g.i;
                        }
                    }
                }
                continue;
            }
            for (int k = 0; k < q; ++k) {
                const double wa = a0 + (a1 - a0) * x[2 * k], fa = (a1 - a0) * x[2 * k + 1] * tent(wa, ca, ha);
                for (int l = 0; l < q; ++l) {
                    const double wb = b0 + (b1 - b0) * x[2 * l];
                    const double r = sqrt(wa * wa + wb * wb + d * d);
                    const Cplx g = rg(G, r);
                    const double f = fa * (b1 - b0) * x[2 * l + 1] * tent(wb, cb, hb) / r;
                    ar += f * g.r;
                    ai += f * g.i;
                }
            }
        }
    return Cplx{ar, ai};
}

// Face pair of perpendicular faces: the integral of tent_c(w) g(sqrt(w^2 +
// u^2 + v^2)) over w (tent centred at cc, half width hc, along the shared
// axis), u in [u0, u1] and v in [v0, v1].  Regular pieces use a q^3 rule; with
// the origin in the closed box the ranges are also split at the origin and a
// piece with the origin at a corner is integrated over three Duffy pyramids
// with their apex there.
__device__ Cplx perp_pair(const GreenCfg& G, double cc, double hc, double u0, double u1, double v0, double v1, int q,
                          int qs, const double* gl)
{
    const double h = fmax(hc, fmax(u1 - u0, v1 - v0));
    const bool sing = cc - hc <= SNAP * h && cc + hc >= -SNAP * h && u0 <= SNAP * h && u1 >= -SNAP * h &&
                      v0 <= SNAP * h && v1 >= -SNAP * h;
    double bw[5], bu[4], bv[4];
    const int nw = breaks(cc - hc, cc, cc + hc, sing, h, bw);
    const int nu = breaks(u0, NAN, u1, sing, h, bu), nv = breaks(v0, NAN, v1, sing, h, bv);
    const double* x = gl_tab(gl, q);
    const double* xs = gl_tab(gl, qs);
    double ar = 0, ai = 0;
    for (int i = 0; i + 1 < nw; ++i)
        for (int j = 0; j + 1 < nu; ++j)
            for (int k = 0; k + 1 < nv; ++k) {
                const double w0 = bw[i], w1 = bw[i + 1], a0 = bu[j], a1 = bu[j + 1], b0 = bv[k], b1 = bv[k + 1];
                if (sing && (w0 == 0.0 || w1 == 0.0) && (a0 == 0.0 || a1 == 0.0) && (b0 == 0.0 || b1 == 0.0)) {
                    const double sw = w0 == 0.0 ? 1.0 : -1.0;
                    const double L[3] = {w1 - w0, a1 - a0, b1 - b0};
                    for (int ax = 0; ax < 3; ++ax) {
                        const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
                        for (int e = 0; e < qs; ++e) {
                            const double t = xs[2 * e], wt = xs[2 * e + 1];
                            for (int f1 = 0; f1 < qs; ++f1) {
                                const double s1 = L[o1] * xs[2 * f1], w1_ = L[o1] * xs[2 * f1 + 1];
                                for (int f2 = 0; f2 < qs; ++f2) {
                                    const double s2 = L[o2] * xs[2 * f2], w2_ = L[o2] * xs[2 * f2 + 1];
                                    double P[3];
                                    P[ax] = t * L[ax];
                                    P[o1] = t * s1;
                                    P[o2] = t * s2;
                                    const double rho = sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
                                    const Cplx g = rg(G, rho);
                                    const double f = wt * w1_ * w2_ * t * t * L[ax] / rho * tent(sw * P[0], cc, hc);
                                    ar += f * g.r;
                                    ai += f * g.i;
                                }
                            }
                        }
                    }
                    continue;
                }
                for (int e = 0; e < q; ++e) {
                    const double w = w0 + (w1 - w0) * x[2 * e], fw = (w1 - w0) * x[2 * e + 1] * tent(w, cc, hc);
                    for (int l = 0; l < q; ++l) {
                        const double u = a0 + (a1 - a0) * x[2 * l], fu = fw * (a1 - a0) * x[2 * l + 1];
                        for (int m = 0; m < q; ++m) {
                            const double v = b0 + (b1 - b0) * x[2 * m];
                            const double r = sqrt(w * w + u * u + v * v);
                            const Cplx g = rg(G, r);
                            const double f = fu * (b1 - b0) * x[2 * m + 1] / r;
                            ar += f * g.r;
                            ai += f * g.i;
                        }
                    }
                }
            }
    return Cplx{ar, ai};
}

// integral of g over target face (normal nt, side st: 0 lower, 1 upper) of
// the cell at separation sep times source face (ns, ss) of the cell at the
// origin (the reference's egoSrfAdp! integral times the face areas)
__device__ Cplx face_pair(const GreenCfg& G, const double sep[3], int nt, int st, int ns, int ss, int q, int qs,
                          const double* gl)
{
    const double* h = G.scl;
    if (nt == ns) {
        const int a = (nt + 1) % 3, b = (nt + 2) % 3;
        return par_pair(G, sep[nt] + (st - ss) * h[nt], sep[a], h[a], sep[b], h[b], q, qs, gl);
    }
    const int c = 3 - nt - ns;
    return perp_pair(G, sep[c], h[c], sep[nt] + (st - 1) * h[nt], sep[nt] + st * h[nt], sep[ns] - ss * h[ns],
                     sep[ns] + (1 - ss) * h[ns], q, qs, gl);
}

// corVal / mask tables of egoFunSng! (src/glaGen.jl:350-525): entry
// [o1 + 2 o2 + 4 o3][target face][source face] (faces 2*normal + side, the
// cubFac order yzL yzU xzL xzU xyL xyU) is 0 for a regular face pair, else
// 1 + the index into [wS(3) wE(9) wV(6)]
#define S1 1
#define S2 2
#define S3 3
#define E1 4
#define E2 5
#define E3 6
#define E4 7
#define E5 8
#define E6 9
#define E7 10
#define E8 11
#define E9 12
#define V1 13
#define V2 14
#define V3 15
#define V4 16
#define V5 17
#define V6 18
#define __ 0
__constant__ unsigned char COR[8][6][6] = {
    {{S1, __, E7, E7, E8, E8}, {__, S1, E7, E7, E8, E8}, {E7, E7, S2, __, E9, E9},
     {E7, E7, __, S2, E9, E9}, {E8, E8, E9, E9, S3, __}, {E8, E8, E9, E9, __, S3}},
    {{__, S1, E7, E7, E8, E8}, {__, __, __, __, __, __}, {__, E7, E3, __, V6, V6},
     {__, E7, __, E3, V6, V6}, {__, E8, V6, V6, E5, __}, {__, E8, V6, V6, __, E5}},
    {{E1, __, __, E7, V5, V5}, {__, E1, __, E7, V5, V5}, {E7, E7, __, S2, E9, E9},
     {__, __, __, __, __, __}, {V5, V5, __, E9, E6, __}, {V5, V5, __, E9, __, E6}},
    {{__, E1, __, E7, V5, V5}, {__, __, __, __, __, __}, {__, E7, __, E3, V6, V6},
     {__, __, __, __, __, __}, {__, V5, __, V6, V3, __}, {__, V5, __, V6, __, V3}},
    {{E2, __, V4, V4, __, E8}, {__, E2, V4, V4, __, E8}, {V4, V4, E4, __, __, E9},
     {V4, V4, __, E4, __, E9}, {E8, E8, E9, E9, __, S3}, {__, __, __, __, __, __}},
    {{__, E2, V4, V4, __, E8}, {__, __, __, __, __, __}, {__, V4, V2, __, __, V6},
     {__, V4, __, V2, __, V6}, {__, E8, V6, V6, __, E5}, {__, __, __, __, __, __}},
    {{V1, __, __, V4, __, V5}, {__, V1, __, V4, __, V5}, {V4, V4, __, E4, __, E9},
     {__, __, __, __, __, __}, {V5, V5, __, E9, __, E6}, {__, __, __, __, __, __}},
    {{__, V1, __, V4, __, V5}, {__, __, __, __, __, __}, {__, V4, __, V2, __, V6},
     {__, __, __, __, __, __}, {__, V5, __, V6, __, V3}, {__, __, __, __, __, __}}};
#undef S1
#undef S2
#undef S3
#undef E1
#undef E2
#undef E3
#undef E4
#undef E5
#undef E6
#undef E7
#undef E8
#undef E9
#undef V1
#undef V2
#undef V3
#undef V4
#undef V5
#undef V6
#undef __

// order of the regular rules: by the separation in cells (Chebyshev) and at
// least the order the cell's phase span needs
__device__ __forceinline__ int reg_order(const GreenCfg& G, double D)
{
    const int q = D < 2.5 ? 12 : (D < 4.5 ? 10 : (D < 8.5 ? 8 : (D < 16.5 ? 6 : (D < 48.5 ? 5 : 4))));
    return q > G.qphase ? q : G.qphase;
}

// six upper-triangle components at separation sep (the srfSum! combination,
// src/glaGen.jl:531-555, written per normal pair): A[i][j] sums target
// normal i, source normal j with side signs (lower +, upper -), negated for
// i != j, divided by the cell volume; xx = A_yy + A_zz, ..., xy = A_xy, ...
// Separations of at most one cell per axis on the cell grid take the
// singular values w18 (already divided by the volume) for the face pairs the
// tables mark, at the absolute offset, with the flips of negative axes.
__device__ void green_at(const GreenCfg& G, const double sep[3], const double* __restrict__ gl,
                         const double2* __restrict__ w18, double out[12])
{
    const double* h = G.scl;
    bool near = true;
    int o[3];
    double f[3], D = 0;
    for (int d = 0; d < 3; ++d) {
        const double t = sep[d] / h[d];
        const double r = rint(t);
        if (fabs(t - r) > 1e-9 || fabs(r) > 1.0) near = false;
        o[d] = (int)fabs(r);
        f[d] = r < 0 ? -1.0 : 1.0;
        D = fmax(D, fabs(t));
    }
    const int q = reg_order(G, D);
    const int qs = G.qphase > 16 ? G.qphase : 16;
    const double vol = h[0] * h[1] * h[2];
    double A[3][3][2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j][0] = A[i][j][1] = 0;
    if (near) {
        const double as[3] = {o[0] * h[0], o[1] * h[1], o[2] * h[2]};
        const int key = o[0] + 2 * o[1] + 4 * o[2];
        for (int nt = 0; nt < 3; ++nt)
            for (int ns = nt; ns < 3; ++ns)
                for (int st = 0; st < 2; ++st)
                    for (int ss = 0; ss < 2; ++ss) {
                        const int code = COR[key][2 * nt + st][2 * ns + ss];
                        Cplx v;
                        if (code) {
                            v = Cplx{w18[code - 1].x, w18[code - 1].y};
                        } else {
                            v = face_pair(G, as, nt, st, ns, ss, q, qs, gl);
                            v.r /= vol;
                            v.i /= vol;
                        }
                        const double sg = (st == ss ? 1.0 : -1.0) * (nt == ns ? 1.0 : -1.0);
                        A[nt][ns][0] += sg * v.r;
                        A[nt][ns][1] += sg * v.i;
                    }
    } else {
        for (int n = 0; n < 3; ++n) {
            // 2 I(d) - I(d + h) - I(d - h) along the normal
            const Cplx i0 = face_pair(G, sep, n, 0, n, 0, q, qs, gl);
            const Cplx ip = face_pair(G, sep, n, 1, n, 0, q, qs, gl);
            const Cplx im = face_pair(G, sep, n, 0, n, 1, q, qs, gl);
            A[n][n][0] = (2 * i0.r - ip.r - im.r) / vol;
            A[n][n][1] = (2 * i0.i - ip.i - im.i) / vol;
        }
        for (int nt = 0; nt < 3; ++nt)
            for (int ns = nt + 1; ns < 3; ++ns)
                for (int st = 0; st < 2; ++st)
                    for (int ss = 0; ss < 2; ++ss) {
                        const Cplx v = face_pair(G, sep, nt, st, ns, ss, q, qs, gl);
                        const double sg = -(st == ss ? 1.0 : -1.0) / vol;
                        A[nt][ns][0] += sg * v.r;
                        A[nt][ns][1] += sg * v.i;
                    }
    }
    out[0] = A[1][1][0] + A[2][2][0];  out[1] = A[1][1][1] + A[2][2][1];
    out[2] = A[0][0][0] + A[2][2][0];  out[3] = A[0][0][1] + A[2][2][1];
    out[4] = A[0][0][0] + A[1][1][0];  out[5] = A[0][0][1] + A[1][1][1];
    const double fx = near ? f[0] * f[1] : 1.0, fy = near ? f[0] * f[2] : 1.0, fz = near ? f[1] * f[2] : 1.0;
    out[6] = fx * A[0][1][0];  out[7] = fx * A[0][1][1];
    out[8] = fy * A[0][2][0];  out[9] = fy * A[0][2][1];
    out[10] = fz * A[1][2][0]; out[11] = fz * A[1][2][1];
    if (sep[0] == 0.0 && sep[1] == 0.0 && sep[2] == 0.0) {
        // identity term -1/zeta^2 (src/glaGen.jl:128-130)
        const double z2r = G.zr * G.zr - G.zi * G.zi, z2i = 2 * G.zr * G.zi;
        const double inv = 1.0 / (z2r * z2r + z2i * z2i);
        for (int c = 0; c < 3; ++c) {
            out[2 * c] -= z2r * inv;
            out[2 * c + 1] -= -z2i * inv;
        }
    }
}

// self Toeplitz tensor over offsets >= 0: toe[(p*6 + comp)] (complex)
__global__ void k_toeplitz(double2* __restrict__ toe, int N1, int N2, int N3, GreenCfg G,
                           const double* __restrict__ gl, const double2* __restrict__ w18)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)N1 * N2 * N3;
    if (p >= n) return;
    const int a = (int)(p % N1), b = (int)((p / N1) % N2), c = (int)(p / ((int64_t)N1 * N2));
    const double sep[3] = {a * G.scl[0], b * G.scl[1], c * G.scl[2]};
    double o[12];
    green_at(G, sep, gl, w18, o);
    for (int k = 0; k < 6; ++k) toe[p * 6 + k] = make_double2(o[2 * k], o[2 * k + 1]);
}

// self circulant from the Toeplitz tensor on an L grid (L >= 2N - 1 per
// axis): c[n] = toe[n] for n < N, zero for N <= n <= L - N, and the flips
// f_i f_j of toe[L - n] beyond (L = 2N is the reference's egoToeCrc!).
// Output layout (L1,L2,L3, comp6).
__global__ void k_embed_self(double2* __restrict__ crc, const double2* __restrict__ toe, int N1, int N2, int N3,
                             int L1, int L2, int L3)
{
    const int64_t Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Lt) return;
    const int n1 = (int)(p % L1), n2 = (int)((p / L1) % L2), n3 = (int)(p / ((int64_t)L1 * L2));
    if ((n1 >= N1 && n1 <= L1 - N1) || (n2 >= N2 && n2 <= L2 - N2) || (n3 >= N3 && n3 <= L3 - N3)) {
        for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(0, 0);
        return;
    }
    const int q1 = n1 < N1 ? n1 : L1 - n1, q2 = n2 < N2 ? n2 : L2 - n2, q3 = n3 < N3 ? n3 : L3 - n3;
    const double f1 = n1 < N1 ? 1 : -1, f2 = n2 < N2 ? 1 : -1, f3 = n3 < N3 ? 1 : -1;
    const double2* s = toe + ((int64_t)q1 + (int64_t)N1 * (q2 + (int64_t)N2 * q3)) * 6;
    const double sg[6] = {1, 1, 1, f1 * f2, f1 * f3, f2 * f3};
    for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(sg[k] * s[k].x, sg[k] * s[k].y);
}

// ---- partitioned operators: target and source cells of different sizes ----
// (the volumes never touch: every face pair is regular)

// overlap length at difference w of an interval of length ha and one of
// length hb whose centres differ by c: a trapezoid on [c - (ha+hb)/2,
// c + (ha+hb)/2] with a flat top min(ha, hb) over |w - c| <= |ha - hb| / 2
__device__ __forceinline__ double trap(double w, double c, double ha, double hb)
{
    return fmax(0.0, fmin(fmin(ha, hb), 0.5 * (ha + hb) - fabs(w - c)));
}

// sorted distinct breakpoints of a trapezoid's support (ends, both kinks and
// the origin when inside, where 1/r is steepest); returns their count
__device__ int trap_breaks(double c, double ha, double hb, double* b)
{
    const double hw = 0.5 * (ha + hb), kw = 0.5 * fabs(ha - hb), h = fmax(ha, hb);
    double v[5] = {c - hw, c - kw, c + kw, c + hw, 0.0};
    const int n = (c - hw < 0.0 && c + hw > 0.0) ? 5 : 4;
    for (int k = 0; k < n; ++k)
        if (fabs(v[k]) < SNAP * h) v[k] = 0.0;
    for (int k = 1; k < n; ++k)
        for (int j = k; j > 0 && v[j] < v[j - 1]; --j) {
            const double t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
    int m = 0;
    for (int k = 0; k < n; ++k)
        if (m == 0 || v[k] > b[m - 1] + SNAP * h) b[m++] = v[k];
    return m;
}

// breakpoints of a plain interval [lo, hi] split at the origin when inside
__device__ int span_breaks(double lo, double hi, double* b)
{
    int m = 0;
    b[m++] = lo;
    if (lo < 0.0 && hi > 0.0) b[m++] = 0.0;
    b[m++] = hi;
    return m;
}

// parallel faces at normal distance d: the integral over (wa, wb) of
// g(sqrt(wa^2 + wb^2 + d^2)) trap_a(wa) trap_b(wb)
__device__ Cplx par_mixed(const GreenCfg& G, double d, double ca, double ta, double sa, double cb, double tb,
                          double sb, int q, const double* gl)
{
    double ba[6], bb[6];
    const int na = trap_breaks(ca, ta, sa, ba), nb = trap_breaks(cb, tb, sb, bb);
    const double* x = gl_tab(gl, q);
    double ar = 0, ai = 0;
    for (int i = 0; i + 1 < na; ++i)
        for (int j = 0; j + 1 < nb; ++j) {
            const double a0 = ba[i], a1 = ba[i + 1], b0 = bb[j], b1 = bb[j + 1];
            for (int k = 0; k < q; ++k) {
                const double wa = a0 + (a1 - a0) * x[2 * k], fa = (a1 - a0) * x[2 * k + 1] * trap(wa, ca, ta, sa);
                for (int l = 0; l < q; ++l) {
                    const double wb = b0 + (b1 - b0) * x[2 * l];
                    const double r = sqrt(wa * wa + wb * wb + d * d);
                    const Cplx g = rg(G, r);
                    const double f = fa * (b1 - b0) * x[2 * l + 1] * trap(wb, cb, tb, sb) / r;
                    ar += f * g.r;
                    ai += f * g.i;
                }
            }
        }
    return Cplx{ar, ai};
}

// perpendicular faces: the integral of trap_c(w) g(sqrt(w^2 + u^2 + v^2))
// over w (shared axis), u in [u0, u1] and v in [v0, v1]
__device__ Cplx perp_mixed(const GreenCfg& G, double cc, double tc, double sc, double u0, double u1, double v0,
                           double v1, int q, const double* gl)
{
    double bw[6], bu[3], bv[3];
    const int nw = trap_breaks(cc, tc, sc, bw), nu = span_breaks(u0, u1, bu), nv = span_breaks(v0, v1, bv);
    const double* x = gl_tab(gl, q);
    double ar = 0, ai = 0;
    for (int i = 0; i + 1 < nw; ++i)
        for (int j = 0; j + 1 < nu; ++j)
            for (int k = 0; k + 1 < nv; ++k) {
                const double w0 = bw[i], w1 = bw[i + 1], a0 = bu[j], a1 = bu[j + 1], b0 = bv[k], b1 = bv[k + 1];
                for (int e = 0; e < q; ++e) {
                    const double w = w0 + (w1 - w0) * x[2 * e], fw = (w1 - w0) * x[2 * e + 1] * trap(w, cc, tc, sc);
                    for (int l = 0; l < q; ++l) {
                        const double u = a0 + (a1 - a0) * x[2 * l], fu = fw * (a1 - a0) * x[2 * l + 1];
                        for (int m = 0; m < q; ++m) {
                            const double v = b0 + (b1 - b0) * x[2 * m];
                            const double r = sqrt(w * w + u * u + v * v);
                            const Cplx g = rg(G, r);
                            const double f = fu * (b1 - b0) * x[2 * m + 1] / r;
                            ar += f * g.r;
                            ai += f * g.i;
                        }
                    }
                }
            }
    return Cplx{ar, ai};
}

// the six components between a target cell (sizes G.scl) and a source cell
// (G.sclS) at separation sep, by the srfSum! combination as green_at forms
// it.  Each face pair carries the reference's srfScl factor
// (src/glaInt.jl:284-305) over the two face areas that its unit-square
// parametrisation divides out: srfScl indexes its table with the target
// face where the source face belongs and back, so the factor is
// area_S(normal of ft) / (A_T(ft) A_S(fs) hT[normal of fs]), with area_S a
// face area in source-cell sizes -- 1 / V_T for cells of equal shape.
__device__ void green_mixed(const GreenCfg& G, const double sep[3], const double* __restrict__ gl, double out[12])
{
    const double* hT = G.scl;
    const double* hS = G.sclS;
    double D = 0;
    for (int d = 0; d < 3; ++d) D = fmax(D, fabs(sep[d]) / (0.5 * (hT[d] + hS[d])));
    const int q = reg_order(G, D);
    double A[3][3][2];
    for (int nt = 0; nt < 3; ++nt)
        for (int ns = nt; ns < 3; ++ns) {
            const int at = (nt + 1) % 3, bt = (nt + 2) % 3, as = (ns + 1) % 3, bs = (ns + 2) % 3;
            const double fac = (hS[at] * hS[bt]) / ((hT[at] * hT[bt]) * (hS[as] * hS[bs]) * hT[ns]);
            double acc_r = 0, acc_i = 0;
            for (int st = 0; st < 2; ++st)
                for (int ss = 0; ss < 2; ++ss) {
                    Cplx v;
                    if (nt == ns) {
                        const double d = sep[nt] + (st - 0.5) * hT[nt] - (ss - 0.5) * hS[nt];
                        v = par_mixed(G, d, sep[at], hT[at], hS[at], sep[bt], hT[bt], hS[bt], q, gl);
                    } else {
                        const int c = 3 - nt - ns;
                        const double ut = sep[nt] + (st - 0.5) * hT[nt], vs = sep[ns] - (ss - 0.5) * hS[ns];
                        v = perp_mixed(G, sep[c], hT[c], hS[c], ut - 0.5 * hS[nt], ut + 0.5 * hS[nt],
                                       vs - 0.5 * hT[ns], vs + 0.5 * hT[ns], q, gl);
                    }
                    const double sg = (st == ss ? 1.0 : -1.0) * (nt == ns ? 1.0 : -1.0);
                    acc_r += sg * v.r;
                    acc_i += sg * v.i;
                }
            A[nt][ns][0] = acc_r * fac;
            A[nt][ns][1] = acc_i * fac;
        }
    out[0] = A[1][1][0] + A[2][2][0];  out[1] = A[1][1][1] + A[2][2][1];
    out[2] = A[0][0][0] + A[2][2][0];  out[3] = A[0][0][1] + A[2][2][1];
    out[4] = A[0][0][0] + A[1][1][0];  out[5] = A[0][0][1] + A[1][1][1];
    out[6] = A[0][1][0];  out[7] = A[0][1][1];
    out[8] = A[0][2][0];  out[9] = A[0][2][1];
    out[10] = A[1][2][0]; out[11] = A[1][2][1];
}

// circulant of one (source, target) partition pair: separations on the
// lattice of step M (the coarser cell size per axis, the grid step of both
// partitions, src/glaMemSup.jl:64-89 with sepGrd, src/glaGen.jl:603-629);
// zero for T <= n <= L - S as k_crc_ext
__global__ void k_crc_prt(double2* __restrict__ crc, int T1, int T2, int T3, int S1, int S2, int S3, int L1, int L2,
                          int L3, double d1, double d2, double d3, double m1, double m2, double m3, GreenCfg G,
                          const double* __restrict__ gl)
{
    const int64_t Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Lt) return;
    const int n1 = (int)(p % L1), n2 = (int)((p / L1) % L2), n3 = (int)(p / ((int64_t)L1 * L2));
    if ((n1 >= T1 && n1 <= L1 - S1) || (n2 >= T2 && n2 <= L2 - S2) || (n3 >= T3 && n3 <= L3 - S3)) {
        for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(0, 0);
        return;
    }
    const double sep[3] = {d1 + (n1 < T1 ? n1 : n1 - L1) * m1, d2 + (n2 < T2 ? n2 : n2 - L2) * m2,
                           d3 + (n3 < T3 ? n3 : n3 - L3) * m3};
    double o[12];
    green_mixed(G, sep, gl, o);
    for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(o[2 * k], o[2 * k + 1]);
}

// external circulant over target-minus-source separations
// (zero for T <= n <= L - S on an axis; L = S + T is the reference's grid)
__global__ void k_crc_ext(double2* __restrict__ crc, int T1, int T2, int T3, int S1, int S2, int S3, int L1, int L2,
                          int L3, double d1, double d2, double d3, GreenCfg G, const double* __restrict__ gl,
                          const double2* __restrict__ w18)
{
    const int64_t Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Lt) return;
    const int n1 = (int)(p % L1), n2 = (int)((p / L1) % L2), n3 = (int)(p / ((int64_t)L1 * L2));
    if ((n1 >= T1 && n1 <= L1 - S1) || (n2 >= T2 && n2 <= L2 - S2) || (n3 >= T3 && n3 <= L3 - S3)) {
        for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(0, 0);
        return;
    }
    const double sep[3] = {d1 + (n1 < T1 ? n1 : n1 - L1) * G.scl[0], d2 + (n2 < T2 ? n2 : n2 - L2) * G.scl[1],
                           d3 + (n3 < T3 ? n3 : n3 - L3) * G.scl[2]};
    double o[12];
    green_at(G, sep, gl, w18, o);
    for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(o[2 * k], o[2 * k + 1]);
}

// branch extraction with truncation: fur[((((b*P + pair)*6 + c)*U3 + j3)*U2 +
// j2)*U1 + j1] (partition pair `pair` of P) from the split-transformed grid W
// (L1,L2,L3, comp6), parity-major per axis; a branch shorter than its stored
// extent repeats its last entry
__global__ void k_gather(double2* __restrict__ fur, const double2* __restrict__ W, int B1, int B2, int B3, int U1,
                         int U2, int U3, int pair, int P)
{
    const int64_t Ut = (int64_t)U1 * U2 * U3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Ut * 48) return;
    const int64_t j = p % Ut;
    const int bc = (int)(p / Ut), c = bc % 6, b = bc / 6;
    const int j1 = (int)(j % U1), j2 = (int)((j / U1) % U2), j3 = (int)(j / ((int64_t)U1 * U2));
    const int k1 = j1 < B1 ? j1 : B1 - 1, k2 = j2 < B2 ? j2 : B2 - 1, k3 = j3 < B3 ? j3 : B3 - 1;
    const int m1 = ((b >> 2) & 1) * B1 + k1, m2 = ((b >> 1) & 1) * B2 + k2, m3 = (b & 1) * B3 + k3;
    const int64_t L1 = 2 * B1, L2 = 2 * B2, Lt = L1 * L2 * 2 * B3;
    fur[(((int64_t)b * P + pair) * 6 + c) * Ut + j] = W[c * Lt + m1 + L1 * (m2 + L2 * (int64_t)m3)];
}

__global__ void k_count_nonfinite(const double2* __restrict__ v, int64_t n, int* __restrict__ cnt)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n && !(isfinite(v[p].x) && isfinite(v[p].y))) atomicAdd(cnt, 1);
}

void gl_rule(int n, double* x, double* w)
{
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            const double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * pp * pp);
    }
}

double rat(int64_t n, int64_t d) { return (double)n / (double)d; }

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// ---- Fourier data at branch lengths without a compiled transform ----

// the full spectrum (L1,L2,L3, 6) of partition pair `pair` of P from the
// stored branches in the reference layout (the mirror rules of mulBrnHst!,
// src/glaActSup.jl:691-753, on symmetric axes; frequency k = 2j + p)
__global__ void k_expand(double2* __restrict__ F, const double2* __restrict__ fur, AxisMap a1, AxisMap a2,
                         AxisMap a3, int pair, int P)
{
    const int L1 = 2 * a1.B, L2 = 2 * a2.B, L3 = 2 * a3.B;
    const int64_t Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Lt * 6) return;
    const int64_t q = p % Lt;
    const int c = (int)(p / Lt);
    const int k1 = (int)(q % L1), k2 = (int)((q / L1) % L2), k3 = (int)(q / ((int64_t)L1 * L2));
    const int p1 = k1 & 1, p2 = k2 & 1, p3 = k3 & 1;
    int j1, j2, j3, s1, s2, s3;
    axis_lookup(a1, p1, k1 >> 1, j1, s1);
    axis_lookup(a2, p2, k2 >> 1, j2, s2);
    axis_lookup(a3, p3, k3 >> 1, j3, s3);
    const int b = (p1 << 2) | (p2 << 1) | p3;
    const int64_t Ut = (int64_t)a1.U * a2.U * a3.U;
    const double2 v = fur[(((int64_t)b * P + pair) * 6 + c) * Ut + ((int64_t)j3 * a2.U + j2) * a1.U + j1];
    const double sg = c == 3 ? s1 * s2 : (c == 4 ? s1 * s3 : (c == 5 ? s2 * s3 : 1));
    F[p] = make_double2(sg * v.x, sg * v.y);
}

// one axis of an inverse DFT of any length by direct summation in FP64:
// out[.. m ..] = sum_k in[.. k ..] e^{2 pi i k m / L}, the phase read from the
// table of the L roots at the exact index k m mod L; layout (L1,L2,L3, 6)
__global__ void k_idft_axis(double2* __restrict__ out, const double2* __restrict__ in, int L1, int L2, int L3,
                            int axis, const double2* __restrict__ roots)
{
    const int64_t Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Lt * 6) return;
    const int64_t q = p % Lt;
    const int n1 = (int)(q % L1), n2 = (int)((q / L1) % L2);
    const int m = axis == 0 ? n1 : (axis == 1 ? n2 : (int)(q / ((int64_t)L1 * L2)));
    const int L = axis == 0 ? L1 : (axis == 1 ? L2 : L3);
    const int64_t st = axis == 0 ? 1 : (axis == 1 ? L1 : (int64_t)L1 * L2);
    const double2* src = in + (p - (int64_t)m * st);
    double ar = 0, ai = 0;
    int idx = 0;
    for (int k = 0; k < L; ++k) {
        const double2 v = src[(int64_t)k * st];
        const double2 w = roots[idx];
        ar += v.x * w.x - v.y * w.y;
        ai += v.x * w.y + v.y * w.x;
        idx += m;
        if (idx >= L) idx -= L;
    }
    out[p] = make_double2(ar, ai);
}

// the circulant of an exact grid L = S + T re-embedded on a grid L' >= L:
// separations 0..T-1 at n' = n, -(S-1)..-1 at the top of L'; zero elsewhere;
// scaled by `scale` (the inverse transform's 1 / (L1 L2 L3))
__global__ void k_reembed(double2* __restrict__ out, const double2* __restrict__ in, int L1, int L2, int L3,
                          int M1, int M2, int M3, int T1, int T2, int T3, int S1, int S2, int S3, double scale)
{
    const int64_t Mt = (int64_t)M1 * M2 * M3, Lt = (int64_t)L1 * L2 * L3;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Mt * 6) return;
    const int64_t q = p % Mt;
    const int c = (int)(p / Mt);
    const int m1 = (int)(q % M1), m2 = (int)((q / M1) % M2), m3 = (int)(q / ((int64_t)M1 * M2));
    auto src = [](int m, int M, int L, int T, int S) { return m < T ? m : (m > M - S ? m - M + L : -1); };
    const int n1 = src(m1, M1, L1, T1, S1), n2 = src(m2, M2, L2, T2, S2), n3 = src(m3, M3, L3, T3, S3);
    if (n1 < 0 || n2 < 0 || n3 < 0) {
        out[p] = make_double2(0, 0);
        return;
    }
    const double2 v = in[(int64_t)c * Lt + n1 + (int64_t)L1 * (n2 + (int64_t)L2 * n3)];
    out[p] = make_double2(scale * v.x, scale * v.y);
}

// The forward transform of circulants (six components on the L = 2B grid,
// FP64, layout (L1,L2,L3,6)) through the pass kernels with S = T = L on every
// axis, and the branch extraction of k_gather
struct FwdGrid {
    int B[3] = {0, 0, 0};
    int64_t Lt = 0;
    Geom g{};
    DevTab tb{};
    DevBuf dtab, dw1, dw2;
};

hipError_t fwd_grid_init(FwdGrid& f, const int B[3])
{
    int L[3];
    for (int a = 0; a < 3; ++a) {
        f.B[a] = B[a];
        L[a] = 2 * B[a];
    }
    f.Lt = (int64_t)L[0] * L[1] * L[2];
    Geom& g = f.g;
    for (int a = 0; a < 3; ++a) { g.S[a] = L[a]; g.T[a] = L[a]; g.B[a] = B[a]; g.L[a] = L[a]; }
    g.Y1 = L[1];
    g.Z = L[2];
    g.xs = f.Lt * 3;
    g.ys = f.Lt * 3;
    g.w1s = f.Lt * 3;
    g.w2s = f.Lt * 3;
    g.xp = whole_part(g.S);
    g.yp = whole_part(g.T);
    std::vector<double2> tab;
    int64_t off[3][2];
    for (int a = 0; a < 3; ++a) {
        off[a][0] = (int64_t)tab.size();
        for (int n = 0; n < B[a]; ++n) tab.push_back(make_double2(std::cos(-M_PI * n / B[a]), std::sin(-M_PI * n / B[a])));
        off[a][1] = (int64_t)tab.size();
        for (int n = 0; n < B[a]; ++n)
            tab.push_back(make_double2(std::cos(-2 * M_PI * n / B[a]), std::sin(-2 * M_PI * n / B[a])));
    }
    hipError_t e = hipMalloc(&f.dtab.p, tab.size() * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpy(f.dtab.p, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice);
    for (int a = 0; a < 3; ++a) {
        f.tb.phz[a] = (const double2*)f.dtab.p + off[a][0];
        f.tb.tw[a] = (const double2*)f.dtab.p + off[a][1];
    }
    if (e == hipSuccess) e = hipMalloc(&f.dw1.p, (size_t)f.Lt * 6 * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&f.dw2.p, (size_t)f.Lt * 6 * sizeof(double2));
    return e;
}

// transform circulant crc and extract its branches into block `pair` of P of
// dfur (stored extents d.fur_ext)
hipError_t fwd_grid_run(FwdGrid& f, const void* crc, double2* dfur, const gla_dims& d, int pair, int P)
{
    hipError_t e = run_fwd3<double>(f.g, f.tb, crc, f.dw1.p, f.dw2.p, 2, nullptr);
    if (e != hipSuccess) return e;
    const int64_t Ut = d.fur_ext[0] * d.fur_ext[1] * d.fur_ext[2];
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((Ut * 48 + 255) / 256)), dim3(256), 0, nullptr, dfur,
                       (const double2*)f.dw2.p, f.B[0], f.B[1], f.B[2], (int)d.fur_ext[0], (int)d.fur_ext[1],
                       (int)d.fur_ext[2], pair, P);
    return hipGetLastError();
}

}  // namespace

gla_status setup_fur(const gla_opts& opts, const gla_vol& trg, const gla_vol& src, const gla_dims& d,
                     std::vector<std::complex<double>>& fur, std::string& err)
{
#define SETUP_HIP(call)                                                             \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return GLA_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)
    for (int a = 0; a < 3; ++a)
        if (!len_supported((int)d.dev_brn[a])) {
            err = "branch length " + std::to_string(d.dev_brn[a]) + " has no compiled transform";
            return GLA_ERR_UNSUPPORTED;
        }
    GreenCfg G{};
    for (int a = 0; a < 3; ++a) {
        G.scl[a] = rat(trg.scl_num[a], trg.scl_den[a]);
        G.sclS[a] = rat(src.scl_num[a], src.scl_den[a]);
    }
    const int Ps = (int)(d.src_div[0] * d.src_div[1] * d.src_div[2]);
    const int Pt = (int)(d.trg_div[0] * d.trg_div[1] * d.trg_div[2]);
    const bool parts = Ps * Pt > 1;
    G.zr = opts.frq_phz_re;
    G.zi = opts.frq_phz_im;
    if (G.zr == 0 && G.zi == 0) {
        err = "frqPhz must be nonzero";
        return GLA_ERR_ARG;
    }
    {
        const std::complex<double> z(G.zr, G.zi), c = 1.0 / (4.0 * M_PI * z * z);
        G.kr = 2 * M_PI * G.zr;
        G.ki = 2 * M_PI * G.zi;
        G.cr = c.real();
        G.ci = c.imag();
        // a regular rule integrates the phase factor across a cell to about
        // 1e-15 from 4 + 3 k h nodes per axis
        double hmax = 0;
        for (int a = 0; a < 3; ++a) hmax = std::max(hmax, std::max(G.scl[a], G.sclS[a]));
        G.qphase = std::min(QMAX, (int)std::ceil(4.0 + 3.0 * 2.0 * M_PI * std::abs(z) * hmax));
    }
    if (!d.self_opr) {
        // overlap check of genEgoCrcExt! (src/glaGen.jl:64-79): edges of each
        // volume along every axis
        double sl[3], su[3], tl[3], tu[3];
        for (int a = 0; a < 3; ++a) {
            const double so = rat(src.org_num[a], src.org_den[a]), to = rat(trg.org_num[a], trg.org_den[a]);
            sl[a] = so - 0.5 * G.sclS[a] * src.cel[a];
            su[a] = so + 0.5 * G.sclS[a] * src.cel[a];
            tl[a] = to - 0.5 * G.scl[a] * trg.cel[a];
            tu[a] = to + 0.5 * G.scl[a] * trg.cel[a];
        }
        bool overlap = true, aligned = true, touch = true;
        for (int a = 0; a < 3; ++a) {
            if (!(tl[a] < su[a] - 1e-12 && sl[a] < tu[a] - 1e-12)) overlap = false;
            if (!(tl[a] <= su[a] + 1e-12 && sl[a] <= tu[a] + 1e-12)) touch = false;
            const double off = (tl[a] - sl[a]) / G.scl[a];
            if (std::fabs(off - std::round(off)) > 1e-9) aligned = false;
        }
        // volumes whose cell grids coincide may overlap: the operator is then
        // the self operator of the covering grid restricted to the two volumes
        if (overlap && (!aligned || parts)) {
            err = "Source and target volumes are overlapping.";
            return GLA_ERR_OVERLAP;
        }
        // cells of different sizes in contact need the contact-volume
        // construction of egoFunExtCnt! (src/glaGen.jl:192-230), not built
        if (touch && parts) {
            err = "source and target volumes with different cell sizes touch; partitioned operators are built "
                  "for separated volumes only";
            return GLA_ERR_UNSUPPORTED;
        }
    }
    std::vector<double> gl((size_t)(QMAX * (QMAX + 1)));
    for (int q = 1; q <= QMAX; ++q) {
        std::vector<double> x(q), w(q);
        gl_rule(q, x.data(), w.data());
        for (int i = 0; i < q; ++i) {
            gl[(size_t)((q * (q - 1) / 2 + i) * 2)] = x[i];
            gl[(size_t)((q * (q - 1) / 2 + i) * 2 + 1)] = w[i];
        }
    }
    // transform lengths; B > (S+T)/2 embeds the circulant in a longer grid
    const int B[3] = {(int)d.dev_brn[0], (int)d.dev_brn[1], (int)d.dev_brn[2]};
    const int L[3] = {2 * B[0], 2 * B[1], 2 * B[2]};
    const int64_t Lt = (int64_t)L[0] * L[1] * L[2];
    // singular face-pair values of touching cells (DIRECTFN order intOrd);
    // partitioned volumes never touch
    std::complex<double> w18[18] = {};
    if (!parts) {
        const gla_status st = singular_values(G.scl, G.zr, G.zi, opts.int_ord > 0 ? opts.int_ord : 32, w18, err);
        if (st) return st;
    }
    DevBuf dgl, dcrc, dtoe, dfur, dcnt, dw18;
    SETUP_HIP(hipMalloc(&dgl.p, gl.size() * sizeof(double)));
    SETUP_HIP(hipMemcpy(dgl.p, gl.data(), gl.size() * sizeof(double), hipMemcpyHostToDevice));
    SETUP_HIP(hipMalloc(&dw18.p, sizeof(w18)));
    SETUP_HIP(hipMemcpy(dw18.p, w18, sizeof(w18), hipMemcpyHostToDevice));
    SETUP_HIP(hipMalloc(&dcrc.p, (size_t)Lt * 6 * sizeof(double2)));
    FwdGrid fg;
    SETUP_HIP(fwd_grid_init(fg, B));
    const int64_t Ut = d.fur_ext[0] * d.fur_ext[1] * d.fur_ext[2];
    const int P = Ps * Pt;
    SETUP_HIP(hipMalloc(&dfur.p, (size_t)Ut * 48 * P * sizeof(double2)));
    SETUP_HIP(hipMalloc(&dcnt.p, sizeof(int)));
    SETUP_HIP(hipMemset(dcnt.p, 0, sizeof(int)));
    const int thr = 128;
    // centre of the first cell of each whole volume
    double t0[3], s0[3];
    for (int a = 0; a < 3; ++a) {
        t0[a] = rat(trg.org_num[a], trg.org_den[a]) - G.scl[a] * (trg.cel[a] / 2 - (trg.cel[a] % 2 == 0 ? 0.5 : 0));
        s0[a] = rat(src.org_num[a], src.org_den[a]) - G.sclS[a] * (src.cel[a] / 2 - (src.cel[a] % 2 == 0 ? 0.5 : 0));
    }
    for (int pair = 0; pair < P; ++pair) {
        if (d.self_opr) {
            const int64_t Nt = d.trg_cel[0] * d.trg_cel[1] * d.trg_cel[2];
            SETUP_HIP(hipMalloc(&dtoe.p, (size_t)Nt * 6 * sizeof(double2)));
            hipLaunchKernelGGL(k_toeplitz, dim3((unsigned)((Nt + thr - 1) / thr)), dim3(thr), 0, nullptr,
                               (double2*)dtoe.p, (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2], G,
                               (const double*)dgl.p, (const double2*)dw18.p);
            SETUP_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_embed_self, dim3((unsigned)((Lt + 255) / 256)), dim3(256), 0, nullptr,
                               (double2*)dcrc.p, (const double2*)dtoe.p, (int)d.trg_cel[0], (int)d.trg_cel[1],
                               (int)d.trg_cel[2], L[0], L[1], L[2]);
            SETUP_HIP(hipGetLastError());
        } else if (!parts) {
            // separation of the first target cell centre from the first source one
            hipLaunchKernelGGL(k_crc_ext, dim3((unsigned)((Lt + thr - 1) / thr)), dim3(thr), 0, nullptr,
                               (double2*)dcrc.p, (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2],
                               (int)d.src_cel[0], (int)d.src_cel[1], (int)d.src_cel[2], L[0], L[1], L[2],
                               t0[0] - s0[0], t0[1] - s0[1], t0[2] - s0[2], G, (const double*)dgl.p,
                               (const double2*)dw18.p);
            SETUP_HIP(hipGetLastError());
        } else {
            // pair = s + Ps t: partition s of the source (offsets os, x
            // fastest) and t of the target; cell i of a partition is cell
            // i div + o of its volume, so the pair's first separation is
            // t0 + ot hT - s0 - os hS and its lattice step the coarser size
            const int s = pair % Ps, t = pair / Ps;
            const int os[3] = {(int)(s % d.src_div[0]), (int)((s / d.src_div[0]) % d.src_div[1]),
                               (int)(s / (d.src_div[0] * d.src_div[1]))};
            const int ot[3] = {(int)(t % d.trg_div[0]), (int)((t / d.trg_div[0]) % d.trg_div[1]),
                               (int)(t / (d.trg_div[0] * d.trg_div[1]))};
            double del[3], M[3];
            for (int a = 0; a < 3; ++a) {
                del[a] = (t0[a] + ot[a] * G.scl[a]) - (s0[a] + os[a] * G.sclS[a]);
                M[a] = std::max(G.scl[a], G.sclS[a]);
            }
            hipLaunchKernelGGL(k_crc_prt, dim3((unsigned)((Lt + thr - 1) / thr)), dim3(thr), 0, nullptr,
                               (double2*)dcrc.p, (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2],
                               (int)d.src_cel[0], (int)d.src_cel[1], (int)d.src_cel[2], L[0], L[1], L[2], del[0],
                               del[1], del[2], M[0], M[1], M[2], G, (const double*)dgl.p);
            SETUP_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)((Lt * 6 + 255) / 256)), dim3(256), 0, nullptr,
                           (const double2*)dcrc.p, Lt * 6, (int*)dcnt.p);
        SETUP_HIP(hipGetLastError());
        SETUP_HIP(fwd_grid_run(fg, dcrc.p, (double2*)dfur.p, d, pair, P));
    }
    int bad = 0;
    SETUP_HIP(hipMemcpy(&bad, dcnt.p, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) {
        err = "Computed circulant contains non-numeric values.";
        return GLA_ERR_NONFINITE;
    }
    fur.resize((size_t)(Ut * 48 * P));
    SETUP_HIP(hipMemcpy(fur.data(), dfur.p, fur.size() * sizeof(double2), hipMemcpyDeviceToHost));
    return GLA_OK;
#undef SETUP_HIP
}

gla_status refur(const gla_dims& dx, const gla_dims& df, const std::vector<std::complex<double>>& in,
                 std::vector<std::complex<double>>& out, std::string& err)
{
#define SETUP_HIP(call)                                                             \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return GLA_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)
    const int P = (int)(dx.src_div[0] * dx.src_div[1] * dx.src_div[2] * dx.trg_div[0] * dx.trg_div[1] *
                        dx.trg_div[2]);
    int L[3], M[3], Bf[3];
    AxisMap am[3];
    for (int a = 0; a < 3; ++a) {
        L[a] = 2 * (int)dx.brn[a];
        Bf[a] = (int)df.dev_brn[a];
        M[a] = 2 * Bf[a];
        am[a].B = (int)dx.brn[a];
        am[a].U = (int)dx.fur_ext[a];
        am[a].sym = (dx.sym[a] && dx.fur_ext[a] != dx.brn[a]) ? 1 : 0;
        const int T = (int)dx.trg_cel[a];
        am[a].fe = (T + 1) / 2 + (T % 2 == 0 ? 1 : 0);
        am[a].fo = (T + 1) / 2;
        if (!len_supported(Bf[a])) {
            err = "branch length " + std::to_string(Bf[a]) + " has no compiled transform";
            return GLA_ERR_UNSUPPORTED;
        }
    }
    const int64_t Lt = (int64_t)L[0] * L[1] * L[2], Mt = (int64_t)M[0] * M[1] * M[2];
    const int64_t Uin = dx.fur_ext[0] * dx.fur_ext[1] * dx.fur_ext[2];
    const int64_t Uout = df.fur_ext[0] * df.fur_ext[1] * df.fur_ext[2];
    DevBuf din, dA, dB, droots, dcrc, dfur;
    SETUP_HIP(hipMalloc(&din.p, (size_t)Uin * 48 * P * sizeof(double2)));
    SETUP_HIP(hipMemcpy(din.p, in.data(), (size_t)Uin * 48 * P * sizeof(double2), hipMemcpyHostToDevice));
    SETUP_HIP(hipMalloc(&dA.p, (size_t)Lt * 6 * sizeof(double2)));
    SETUP_HIP(hipMalloc(&dB.p, (size_t)Lt * 6 * sizeof(double2)));
    SETUP_HIP(hipMalloc(&dcrc.p, (size_t)Mt * 6 * sizeof(double2)));
    SETUP_HIP(hipMalloc(&dfur.p, (size_t)Uout * 48 * P * sizeof(double2)));
    // roots e^{2 pi i r / L} of each axis, back to back
    std::vector<double2> roots;
    int roff[3];
    for (int a = 0; a < 3; ++a) {
        roff[a] = (int)roots.size();
        for (int r = 0; r < L[a]; ++r)
            roots.push_back(make_double2(std::cos(2 * M_PI * r / L[a]), std::sin(2 * M_PI * r / L[a])));
    }
    SETUP_HIP(hipMalloc(&droots.p, roots.size() * sizeof(double2)));
    SETUP_HIP(hipMemcpy(droots.p, roots.data(), roots.size() * sizeof(double2), hipMemcpyHostToDevice));
    FwdGrid fg;
    SETUP_HIP(fwd_grid_init(fg, Bf));
    const unsigned gl = (unsigned)((Lt * 6 + 255) / 256), gm = (unsigned)((Mt * 6 + 255) / 256);
    for (int pair = 0; pair < P; ++pair) {
        hipLaunchKernelGGL(k_expand, dim3(gl), dim3(256), 0, nullptr, (double2*)dA.p, (const double2*)din.p, am[0],
                           am[1], am[2], pair, P);
        SETUP_HIP(hipGetLastError());
        void* src = dA.p;
        void* dst = dB.p;
        for (int a = 0; a < 3; ++a) {
            hipLaunchKernelGGL(k_idft_axis, dim3(gl), dim3(256), 0, nullptr, (double2*)dst, (const double2*)src, L[0],
                               L[1], L[2], a, (const double2*)droots.p + roff[a]);
            SETUP_HIP(hipGetLastError());
            std::swap(src, dst);
        }
        hipLaunchKernelGGL(k_reembed, dim3(gm), dim3(256), 0, nullptr, (double2*)dcrc.p, (const double2*)src, L[0],
                           L[1], L[2], M[0], M[1], M[2], (int)dx.trg_cel[0], (int)dx.trg_cel[1], (int)dx.trg_cel[2],
                           (int)dx.src_cel[0], (int)dx.src_cel[1], (int)dx.src_cel[2], 1.0 / (double)Lt);
        SETUP_HIP(hipGetLastError());
        SETUP_HIP(fwd_grid_run(fg, dcrc.p, (double2*)dfur.p, df, pair, P));
    }
    out.resize((size_t)(Uout * 48 * P));
    SETUP_HIP(hipMemcpy(out.data(), dfur.p, out.size() * sizeof(double2), hipMemcpyDeviceToHost));
    return GLA_OK;
#undef SETUP_HIP
}

}  // namespace gla
