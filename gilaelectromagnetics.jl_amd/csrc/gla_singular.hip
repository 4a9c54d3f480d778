// Singular face-pair integrals of touching cells, FP64 on the GPU: the
// reference's DIRECTFN_E reductions of the weakly singular triangle-pair
// integrals (wekSInt / wekEInt / wekVInt and their helpers,
// src/glaIntSup.jl:13-469), summed over the square panels of a cuboid cell as
// wekS / wekE / wekV do (src/glaInt.jl:75-267, point grid wekGrdPts!
// :356-398), plus the closed-form static panel integrals rSrfSlf / rSrfEdgCrn
// / rSrfEdgFlt (src/glaInt.jl:410-488).  The result is the 18 values
// [wS(3), wE(9), wV(6)] divided by the cell volume that egoFunSng!
// substitutes for the face pairs of touching cells (src/glaGen.jl:117-121,
// 350-525).
//
// Work layout: the 108 triangle-pair integrals of one cell shape are split
// into independent quadrature items (outer Gauss-Legendre indices of each
// reduction); one thread computes one item, a second kernel sums each
// integral's items in a fixed order (one workgroup per integral), and the host
// combines the 108 sums with the closed forms.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "gla_setup.hpp"

namespace gla {
namespace {

constexpr double SQ3 = 1.7320508075688772935274463415059;

struct Tri2 {
    double r[6][3];   // two triangles (r0,r1,r2) and (r3,r4,r5)
};

struct Zeta {
    double kr, ki;    // 2 pi zeta
    double cr, ci;    // 1 / (4 pi zeta^2)
};

struct Z {
    double r, i;
};

__device__ __forceinline__ Z zmul(Z a, Z b) { return Z{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

// sclEgo: exp(2 pi i zeta d) / (4 pi zeta^2 d)
__device__ __forceinline__ Z ego(const Zeta& z, double d)
{
    double s, c;
    sincos(z.kr * d, &s, &c);
    const double m = (z.ki == 0.0 ? 1.0 : exp(-z.ki * d)) / d;
    return zmul(Z{m * c, m * s}, Z{z.cr, z.ci});
}

// sclEgoN: the same with the static part removed; ((i/zeta) - pi d)/2 below
// 1e-7 (src/glaInt.jl:51-59)
__device__ __forceinline__ Z ego_n(const Zeta& z, double zr, double zi, double d)
{
    if (d > 1e-7) {
        double s, c;
        sincos(z.kr * d, &s, &c);
        const double m = z.ki == 0.0 ? 1.0 : exp(-z.ki * d);
        const Z e{(m * c - 1.0) / d, m * s / d};
        return zmul(e, Z{z.cr, z.ci});
    }
    // i / zeta = i conj(zeta) / |zeta|^2
    const double n2 = zr * zr + zi * zi;
    return Z{(zi / n2 - M_PI * d) * 0.5, (zr / n2) * 0.5};
}

__device__ __forceinline__ double thf(double a, double b, double p) { return 0.5 * ((b - a) * p + a + b); }

__device__ __forceinline__ void spx(double X[2][3], double e1, double e2, double x1, double x2)
{
    X[0][0] = (SQ3 * (1.0 - e1) - x1) / (2.0 * SQ3);
    X[0][1] = (SQ3 * (1.0 + e1) - x1) / (2.0 * SQ3);
    X[0][2] = x1 / SQ3;
    X[1][0] = (SQ3 * (1.0 - e2) - x2) / (2.0 * SQ3);
    X[1][1] = (SQ3 * (1.0 + e2) - x2) / (2.0 * SQ3);
    X[1][2] = x2 / SQ3;
}

__device__ __forceinline__ double dist(const Tri2& T, const double X[2][3], bool self)
{
    const int o = self ? 0 : 3;
    double s = 0;
    for (int d = 0; d < 3; ++d) {
        const double v = X[0][0] * T.r[0][d] + X[0][1] * T.r[1][d] + X[0][2] * T.r[2][d] -
                         (X[1][0] * T.r[o][d] + X[1][1] * T.r[o + 1][d] + X[1][2] * T.r[o + 2][d]);
        s += v * v;
    }
    return sqrt(s);
}

struct Ctx {
    int n;
    const double* x;   // GL nodes on [-1, 1]
    const double* w;   // GL weights
    Zeta z;
    double zr, zi;
};

// ---- self triangle -------------------------------------------------------
__device__ void psi_s(int idf, double& a, double& b)
{
    if (idf == 1 || idf == 5 || idf == 6) { a = 0.0; b = M_PI / 3.0; }
    else if (idf == 2 || idf == 7) { a = M_PI / 3.0; b = 2.0 * M_PI / 3.0; }
    else { a = 2.0 * M_PI / 3.0; b = M_PI; }
}

__device__ void eta_s(int idf, double th, double& a, double& b)
{
    if (idf == 1 || idf == 2) { a = 0.0; b = 1.0; }
    else if (idf == 3) { const double t = tan(M_PI - th); a = (SQ3 - t) / (SQ3 + t); b = 1.0; }
    else if (idf == 4) { const double t = tan(M_PI - th); a = 0.0; b = (SQ3 - t) / (SQ3 + t); }
    else if (idf == 5) { const double t = tan(th); a = (t - SQ3) / (SQ3 + t); b = 0.0; }
    else if (idf == 6) { const double t = tan(th); a = -1.0; b = (t - SQ3) / (SQ3 + t); }
    else { a = -1.0; b = 0.0; }
}

__device__ __forceinline__ void sub_tri(double l1, double l2, int dir, double& e, double& x)
{
    if (dir == 1) { e = l1; x = l2; }
    else if (dir == 2) { e = 0.5 * (1.0 - l1 - l2 * SQ3); x = 0.5 * (SQ3 + l1 * SQ3 - l2); }
    else { e = 0.5 * (-1.0 - l1 + l2 * SQ3); x = 0.5 * (SQ3 - l1 * SQ3 - l2); }
}

// aS: innermost radial integral of the self reduction
__device__ Z a_s(const Ctx& c, const Tri2& T, double s1, double c1, double tb, double th, int dir)
{
    double e1, x1;
    sub_tri(tb, th * s1, dir, e1, x1);
    Z acc{0, 0};
    for (int n = 0; n < c.n; ++n) {
        const double t = thf(0.0, th, c.x[n]);
        double e2, x2, X[2][3];
        sub_tri(t * c1 + tb, (th - t) * s1, dir, e2, x2);
        spx(X, e1, e2, x1, x2);
        const Z g = ego_n(c.z, c.zr, c.zi, dist(T, X, true));
        acc.r += c.w[n] * t * g.r;
        acc.i += c.w[n] * t * g.i;
    }
    return Z{0.5 * th * acc.r, 0.5 * th * acc.i};
}

// nS
__device__ Z n_s(const Ctx& c, const Tri2& T, int dir, int idf, double t1, double tb)
{
    const double s1 = sin(t1), c1 = cos(t1);
    double lim, pre;
    if (idf == 1 || idf == 5) { lim = (1.0 - tb) / c1; pre = (1.0 - tb) / (2.0 * c1); }
    else if (idf == 2 || idf == 3) { lim = SQ3 * (1.0 - tb) / s1; pre = SQ3 * (1.0 - tb) / (2.0 * s1); }
    else if (idf == 6 || idf == 7) { lim = SQ3 * (1.0 + tb) / s1; pre = SQ3 * (1.0 + tb) / (2.0 * s1); }
    else { lim = -(1.0 + tb) / c1; pre = -(1.0 + tb) / (2.0 * c1); }
    Z acc{0, 0};
    for (int n = 0; n < c.n; ++n) {
        const Z a = a_s(c, T, s1, c1, tb, thf(0.0, lim, c.x[n]), dir);
        acc.r += c.w[n] * a.r;
        acc.i += c.w[n] * a.i;
    }
    return Z{pre * acc.r, pre * acc.i};
}

// ---- edge-adjacent triangles --------------------------------------------
__device__ void psi_e(int idf, double& a, double& b)
{
    if (idf == 1) { a = 0.0; b = M_PI / 3.0; }
    else if (idf == 2 || idf == 3) { a = M_PI / 3.0; b = M_PI / 2.0; }
    else if (idf == 4 || idf == 6) { a = M_PI / 2.0; b = M_PI; }
    else { a = 0.0; b = M_PI / 2.0; }
}

__device__ void eta_e(int idf, double th, double& a, double& b)
{
    double s, c;
    sincos(th, &s, &c);
    if (idf == 1) { a = 0.0; b = atan(s + SQ3 * c); }
    else if (idf == 2) { a = atan(s - SQ3 * c); b = atan(s + SQ3 * c); }
    else if (idf == 3 || idf == 4) { a = 0.0; b = atan(s - SQ3 * c); }
    else if (idf == 5) { a = atan(s + SQ3 * c); b = 0.5 * M_PI; }
    else { a = atan(s - SQ3 * c); b = 0.5 * M_PI; }
}

// aE with spxE!: tb, t1 are the angles called θB, θ1 there
__device__ Z a_e(const Ctx& c, const Tri2& T, double lam, double eta, double sb, double cb, double s1, double c1,
                 int idf)
{
    Z acc{0, 0};
    for (int n = 0; n < c.n; ++n) {
        const double zt = thf(0.0, lam, c.x[n]);
        double X[2][3];
        if (idf == 1) spx(X, eta, zt * cb * c1 - eta, zt * sb, zt * cb * s1);
        else spx(X, -eta, -(zt * cb * c1 - eta), zt * sb, zt * cb * s1);
        const Z g = ego_n(c.z, c.zr, c.zi, dist(T, X, false));
        acc.r += c.w[n] * zt * zt * g.r;
        acc.i += c.w[n] * zt * zt * g.i;
    }
    return Z{0.5 * lam * acc.r, 0.5 * lam * acc.i};
}

// intNE
__device__ Z int_ne(const Ctx& c, const Tri2& T, int n, int k, double g, double tb, double t1, int idf2)
{
    double sb, cb, s1, c1;
    sincos(tb, &sb, &cb);
    sincos(t1, &s1, &c1);
    double eta, lam;
    if (k == 1) { eta = thf(0.0, g, c.x[n]); lam = SQ3 * (1 + eta) / (cb * (s1 + SQ3 * c1)); }
    else if (k == 2) { eta = thf(g, 1.0, c.x[n]); lam = SQ3 * (1.0 - fabs(eta)) / sb; }
    else if (k == 3) { eta = thf(g, 1.0, c.x[n]); lam = SQ3 * (1.0 - fabs(eta)) / (cb * (s1 - SQ3 * c1)); }
    else if (k == 4) { eta = thf(0.0, 1.0, c.x[n]); lam = SQ3 * (1.0 - fabs(eta)) / (cb * (s1 - SQ3 * c1)); }
    else { eta = thf(0.0, 1.0, c.x[n]); lam = SQ3 * (1.0 - eta) / sb; }
    return a_e(c, T, lam, eta, sb, cb, s1, c1, idf2);
}

// nE (third and fourth arguments θB, θ1 there)
__device__ Z n_e(const Ctx& c, const Tri2& T, int idf1, int idf2, double tb, double t1)
{
    Z i1{0, 0}, i2{0, 0};
    if (idf1 == 1 || idf1 == 2) {
        const double s1 = sin(t1), c1 = cos(t1);
        const double g = (s1 + SQ3 * c1 - tan(tb)) / (s1 + SQ3 * c1 + tan(tb));
        for (int n = 0; n < c.n; ++n) {
            const Z a = int_ne(c, T, n, 1, g, tb, t1, idf2), b = int_ne(c, T, n, 2, g, tb, t1, idf2);
            i1.r += c.w[n] * a.r; i1.i += c.w[n] * a.i;
            i2.r += c.w[n] * b.r; i2.i += c.w[n] * b.i;
        }
        return Z{0.5 * i2.r + g * 0.5 * (i1.r - i2.r), 0.5 * i2.i + g * 0.5 * (i1.i - i2.i)};
    }
    if (idf1 == 3) {
        const double g = SQ3 / tan(t1);
        for (int n = 0; n < c.n; ++n) {
            const Z a = int_ne(c, T, n, 1, g, tb, t1, idf2), b = int_ne(c, T, n, 3, g, tb, t1, idf2);
            i1.r += c.w[n] * a.r; i1.i += c.w[n] * a.i;
            i2.r += c.w[n] * b.r; i2.i += c.w[n] * b.i;
        }
        return Z{0.5 * i2.r + 0.5 * g * (i1.r - i2.r), 0.5 * i2.i + 0.5 * g * (i1.i - i2.i)};
    }
    const int k = idf1 == 4 ? 4 : 5;
    for (int n = 0; n < c.n; ++n) {
        const Z a = int_ne(c, T, n, k, 1.0, tb, t1, idf2);
        i1.r += c.w[n] * a.r; i1.i += c.w[n] * a.i;
    }
    return Z{0.5 * i1.r, 0.5 * i1.i};
}

__device__ void store(double2* part, int64_t idx, Z v) { part[idx] = make_double2(v.r, v.i); }

// one thread: item (k, idf, dir, it) of a self-triangle integral
__global__ void k_dfn_s(const Tri2* __restrict__ tris, int n, const double* __restrict__ gx,
                        const double* __restrict__ gw, Zeta z, double zr, double zi, double2* __restrict__ part)
{
    const int ni = 24 * n * n;
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= ni) return;
    const Tri2 T = tris[blockIdx.y];
    const Ctx c{n, gx, gw, z, zr, zi};
    const int it = item % n, dir = (item / n) % 3 + 1, idf = (item / (3 * n)) % 8 + 1, k = item / (24 * n);
    double pa, pb, ea, eb;
    psi_s(idf, pa, pb);
    const double th = thf(pa, pb, gx[k]);
    eta_s(idf, th, ea, eb);
    const Z v = n_s(c, T, dir, idf, th, thf(ea, eb, gx[it]));
    const double f = 0.25 * gw[it] * gw[k] * (pb - pa) * (eb - ea) * sin(th);
    store(part, (int64_t)blockIdx.y * ni + item, Z{f * v.r, f * v.i});
}

// one thread: item (k, idf, it) of an edge-adjacent integral
__global__ void k_dfn_e(const Tri2* __restrict__ tris, int n, const double* __restrict__ gx,
                        const double* __restrict__ gw, Zeta z, double zr, double zi, double2* __restrict__ part)
{
    const int ni = 6 * n * n;
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= ni) return;
    const Tri2 T = tris[blockIdx.y];
    const Ctx c{n, gx, gw, z, zr, zi};
    const int it = item % n, idf = (item / n) % 6 + 1, k = item / (6 * n);
    double pa, pb, ea, eb;
    psi_e(idf, pa, pb);
    const double thb = thf(pa, pb, gx[k]);
    eta_e(idf, thb, ea, eb);
    const double tha = thf(ea, eb, gx[it]);
    const Z p = n_e(c, T, idf, 1, tha, thb), m = n_e(c, T, idf, -1, tha, thb);
    const double f = 0.25 * gw[k] * (pb - pa) * (eb - ea) * gw[it] * cos(tha);
    store(part, (int64_t)blockIdx.y * ni + item, Z{f * (p.r + m.r), f * (p.i + m.i)});
}

// one thread: item (i1, i2, i3) of a vertex-adjacent integral; sng selects
// the regularised kernel (wekVInt's sngMod)
__global__ void k_dfn_v(const Tri2* __restrict__ tris, const int* __restrict__ sng, int n,
                        const double* __restrict__ gx, const double* __restrict__ gw, Zeta z, double zr, double zi,
                        double2* __restrict__ part)
{
    const int ni = n * n * n;
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= ni) return;
    const Tri2 T = tris[blockIdx.y];
    const bool reg = sng[blockIdx.y] != 0;
    const int i1 = item % n, i2 = (item / n) % n, i3 = item / (n * n);
    const double ta = thf(0.0, M_PI / 3.0, gx[i1]);
    double sa, ca;
    sincos(ta, &sa, &ca);
    const double la = 2.0 * SQ3 / (sa + SQ3 * ca);
    const double tb = thf(0.0, M_PI / 3.0, gx[i2]);
    double sb, cb;
    sincos(tb, &sb, &cb);
    const double lb = 2.0 * SQ3 / (sb + SQ3 * cb);
    const double at = atan(lb / la);
    const double tc = thf(0.0, at, gx[i3]), td = thf(at, M_PI / 2.0, gx[i3]);
    double sc, cc, sd, cd;
    sincos(tc, &sc, &cc);
    sincos(td, &sd, &cd);
    Z ic{0, 0}, id{0, 0};
    for (int e = 0; e < n; ++e) {
        const double tx = thf(0.0, la / cc, gx[e]);
        double X[2][3];
        spx(X, tx * cc * ca - 1.0, tx * sc * cb - 1.0, tx * cc * sa, tx * sc * sb);
        const double d = dist(T, X, false);
        const Z g = reg ? ego_n(z, zr, zi, d) : ego(z, d);
        const double f = gw[e] * tx * tx * tx;
        ic.r += f * g.r; ic.i += f * g.i;
    }
    for (int e = 0; e < n; ++e) {
        const double tx = thf(0.0, lb / sd, gx[e]);
        double X[2][3];
        spx(X, tx * cd * ca - 1.0, tx * sd * cb - 1.0, tx * cd * sa, tx * sd * sb);
        const double d = dist(T, X, false);
        const Z g = reg ? ego_n(z, zr, zi, d) : ego(z, d);
        const double f = gw[e] * tx * tx * tx;
        id.r += f * g.r; id.i += f * g.i;
    }
    const double w = gw[i1] * gw[i2] * gw[i3];
    const double a = at * la * sc, b = (0.5 * M_PI - at) * lb * cd;
    store(part, (int64_t)blockIdx.y * ni + item, Z{w * (a * ic.r + b * id.r), w * (a * ic.i + b * id.i)});
}

// sum of each integral's items in a fixed order: one workgroup per integral
__global__ void k_dfn_sum(const double2* __restrict__ part, int64_t ni, double2* __restrict__ out)
{
    __shared__ double sr[256], si[256];
    const double2* p = part + (int64_t)blockIdx.x * ni;
    double ar = 0, ai = 0;
    for (int64_t i = threadIdx.x; i < ni; i += 256) {
        ar += p[i].x;
        ai += p[i].y;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(sr[0], si[0]);
}

// ---- host side -----------------------------------------------------------
using cd = std::complex<double>;

void gl_pm1(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign((size_t)n, 0.0);
    w.assign((size_t)n, 0.0);
    for (int i = 0; i < n; ++i) {
        double zz = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * zz * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (zz * p1 - p2) / (zz * zz - 1);
            const double z1 = zz;
            zz = z1 - p1 / pp;
            if (std::fabs(zz - z1) < 1e-16) break;
        }
        x[(size_t)(n - 1 - i)] = zz;
        w[(size_t)(n - 1 - i)] = 2.0 / ((1 - zz * zz) * pp * pp);
    }
}

// wekGrdPts!: the 3x3x2 point grid of a panel direction; point k (1-based in
// the reference) at ((k-1) % 3, ((k-1)/3) % 3, (k-1)/9) grid steps
void grid_point(int dir, const double s[3], int k, double p[3])
{
    double gx, gy, gz;
    if (dir == 1) { gx = s[0]; gy = s[1]; gz = s[2]; }
    else if (dir == 2) { gx = s[2]; gy = s[0]; gz = s[1]; }
    else { gx = s[1]; gy = s[2]; gz = s[0]; }
    p[0] = ((k - 1) % 3) * gx;
    p[1] = (((k - 1) / 3) % 3) * gy;
    p[2] = ((k - 1) / 9) * gz;
}

struct Plan {
    std::vector<Tri2> s, e, v;
    std::vector<int> vsng;
    // each term: (kind 0 S / 1 E / 2 V, index in that list)
    std::vector<std::vector<std::pair<int, int>>> terms;
};

Tri2 tri(int dir, const double s[3], std::initializer_list<int> pts)
{
    Tri2 t{};
    int k = 0;
    for (int p : pts) grid_point(dir, s, p, t.r[k++]);
    if (k == 3)
        for (int j = 0; j < 3; ++j)
            for (int d = 0; d < 3; ++d) t.r[3 + j][d] = t.r[j][d];
    return t;
}

// a term: the sum of the listed triangle-pair integrals
int add_term(Plan& P, std::vector<std::pair<int, Tri2>> parts, int sng = 0)
{
    std::vector<std::pair<int, int>> t;
    for (auto& pr : parts) {
        if (pr.first == 0) { P.s.push_back(pr.second); t.push_back({0, (int)P.s.size() - 1}); }
        else if (pr.first == 1) { P.e.push_back(pr.second); t.push_back({1, (int)P.e.size() - 1}); }
        else { P.v.push_back(pr.second); P.vsng.push_back(sng); t.push_back({2, (int)P.v.size() - 1}); }
    }
    P.terms.push_back(t);
    return (int)P.terms.size() - 1;
}

// rSrfSlf (src/glaInt.jl:410-432)
cd r_srf_slf(double la, double lb, cd z)
{
    const double a2 = la * la, b2 = lb * lb, r1 = std::sqrt(a2 + b2), r4 = std::sqrt(a2 + 4 * b2);
    const double t1 = 8 * la * a2 + 8 * lb * b2 - 8 * a2 * r1 - 8 * b2 * r1;
    const double t2 = -3 * a2 * lb *
                      (2 * std::log(la) + 2 * std::log(la + lb - r1) + std::log(r1 - lb) - 5 * std::log(lb + r1) -
                       2 * std::log(lb - la + r1) - 2 * std::log(la + 2 * lb - r4) + std::log(r4 - 2 * lb) +
                       2 * std::log(la - 2 * lb + r4) + std::log(2 * lb + r4) + 2 * std::log(2 * lb - la + r4) -
                       2 * std::log(la + 2 * lb + r4));
    const double t3 = 6 * la * b2 *
                      (std::log(64.0) + 4 * std::log(lb) + 2 * std::log(r1 - la) + 3 * std::log(la + r1) -
                       3 * std::log(r4 - la) - 3 * std::log(std::sqrt(a2 * a2 + 5 * a2 * b2 + 4 * b2 * b2) + la * (r1 - la - r4)));
    return (1.0 / (48.0 * M_PI * z * z)) * (t1 + t2 + t3);
}

// rSrfEdgCrn (src/glaInt.jl:438-466): la the shared edge
cd r_srf_edg_crn(double la, double lb, double lc, cd z)
{
    const double a2 = la * la, b2 = lb * lb, c2 = lc * lc, a3 = a2 * la, c3 = c2 * lc;
    const double R = std::sqrt(a2 + b2 + c2), rbc = std::sqrt(b2 + c2);
    const double v =
        8 * lb * lc * rbc - 8 * lb * lc * R - 12 * a3 * std::atan(1.0 / (la * lc / (a2 + b2 - lb * R))) +
        12 * a3 * std::atan(la / lc) - 12 * la * c2 * std::atan(la * lb / (lc * R)) -
        12 * la * b2 * std::atan(la * lc / (lb * R)) - 16 * a3 * std::atan(lb * lc / (la * R)) +
        6 * c3 * std::atanh(lb / rbc) - 6 * lc * (a2 + c2) * std::atanh(lb / R) - 15 * a2 * lc * std::log(a2 + c2) -
        c3 * std::log(a2 + c2) + 2 * c3 * std::log(lc / (lb + rbc)) + 6 * a2 * lc * std::log(R - lb) +
        24 * a2 * lc * std::log(R + lb) + 2 * c3 * std::log(R + lb) +
        6 * la * lb *
            (-2 * la * std::log(a2 + b2) - lc * std::log((b2 + c2) * (R - la)) + 3 * lc * std::log(la + R) +
             la * std::log(R - lc) + 3 * la * std::log(R + lc)) +
        2 * lb * b2 * (std::log((R - lc) / (lc + R)) + std::log(1 + (2 * lc * (lc + rbc)) / b2));
    return (1.0 / (48.0 * M_PI * z * z)) * v;
}

// rSrfEdgFlt (src/glaInt.jl:471-488): la the shared edge; its logarithmic
// part carries 1/(64 pi) without the 1/zeta^2 of the other terms, as the
// reference writes it (the two agree at zeta = 1)
cd r_srf_edg_flt(double la, double lb, cd z)
{
    const double a2 = la * la, b2 = lb * lb, r1 = std::sqrt(a2 + b2), r4 = std::sqrt(a2 + 4 * b2);
    const cd p1 = (1.0 / (12.0 * M_PI * z * z)) * (-la * a2 + 2 * b2 * (3 * lb + r1 - 2 * r4) + a2 * (2 * r1 - r4));
    const double q1 = lb * (-62 * std::log(2.0) - 5 * std::log(-la + r1) + 4 * std::log(8 * b2 * (-la + r1)) -
                            33 * std::log(la + r1) + 17 * std::log(-la + r4) - 24 * std::log(lb * (-la + r4)) +
                            57 * std::log(la + r4));
    const double q2 =
        4 * la *
        (-8 * std::asinh(lb / la) + 6 * std::asinh(2 * lb / la) + 6 * std::atanh(lb / r1) + 12 * std::log(la) -
         13 * std::log(-lb + r1) + std::log((-lb + r1) / la) + std::log(la / (lb + r1)) - 7 * std::log(lb + r1) -
         2 * std::log((lb + r1) / la) - 3 * std::log(-(((lb + r1) * (2 * lb - r4)) / a2)) -
         3 * std::log((-lb + r1) / (-2 * lb + r4)) + 11 * std::log(-2 * lb + r4) -
         3 * std::log((lb + r1) / (2 * lb + r4)) + std::log(2 * lb + r4) + 9 * std::log((2 * lb + r4) / (lb + r1)) -
         2 * std::log(a2 + 2 * lb * (lb - r1)));
    return p1 + (1.0 / (64.0 * M_PI)) * la * lb * (q1 + q2);
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() { if (p) (void)hipFree(p); }
};

}  // namespace

gla_status singular_values(const double s[3], double zr, double zi, int ord, std::complex<double> w18[18],
                           std::string& err)
{
#define SNG_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return GLA_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)
    if (ord < 1 || ord > 64) {
        err = "intOrd must be between 1 and 64";
        return GLA_ERR_ARG;
    }
    Plan P;
    const int S = 0, E = 1, V = 2;
    // wekSDir for dir 3, 2, 1 (terms 0..2)
    for (int dir : {3, 2, 1})
        add_term(P, {{S, tri(dir, s, {1, 2, 5})}, {S, tri(dir, s, {1, 5, 4})},
                     {E, tri(dir, s, {1, 2, 5, 1, 5, 4})}, {E, tri(dir, s, {1, 5, 4, 1, 2, 5})},
                     {S, tri(dir, s, {4, 1, 2})}, {S, tri(dir, s, {4, 2, 5})},
                     {E, tri(dir, s, {4, 1, 2, 4, 2, 5})}, {E, tri(dir, s, {4, 2, 5, 4, 1, 2})}});
    // wekEDir for dir 3, 2, 1: four terms each (terms 3..14), regularised
    // vertex kernels
    for (int dir : {3, 2, 1}) {
        add_term(P, {{E, tri(dir, s, {1, 2, 5, 2, 3, 5})}, {V, tri(dir, s, {1, 2, 5, 3, 6, 5})},
                     {V, tri(dir, s, {1, 5, 4, 2, 3, 5})}, {V, tri(dir, s, {1, 5, 4, 3, 6, 5})}}, 1);
        add_term(P, {{E, tri(dir, s, {1, 2, 5, 2, 11, 5})}, {V, tri(dir, s, {1, 2, 5, 11, 14, 5})},
                     {V, tri(dir, s, {1, 5, 4, 2, 11, 5})}, {V, tri(dir, s, {1, 5, 4, 11, 14, 5})}}, 1);
        add_term(P, {{V, tri(dir, s, {1, 2, 5, 4, 5, 7})}, {V, tri(dir, s, {1, 2, 5, 5, 8, 7})},
                     {E, tri(dir, s, {1, 5, 4, 4, 5, 7})}, {V, tri(dir, s, {1, 5, 4, 5, 8, 7})}}, 1);
        add_term(P, {{V, tri(dir, s, {1, 2, 5, 4, 5, 13})}, {V, tri(dir, s, {1, 2, 5, 5, 14, 13})},
                     {E, tri(dir, s, {1, 5, 4, 4, 5, 13})}, {V, tri(dir, s, {1, 5, 4, 5, 14, 13})}}, 1);
    }
    // wekVDir for dir 3, 2, 1: three terms each (terms 15..23), full kernel
    for (int dir : {3, 2, 1}) {
        add_term(P, {{V, tri(dir, s, {1, 2, 5, 5, 6, 9})}, {V, tri(dir, s, {1, 2, 5, 5, 9, 8})},
                     {V, tri(dir, s, {1, 5, 4, 5, 6, 9})}, {V, tri(dir, s, {1, 5, 4, 5, 9, 8})}}, 0);
        add_term(P, {{V, tri(dir, s, {1, 2, 5, 5, 17, 14})}, {V, tri(dir, s, {1, 2, 5, 5, 8, 17})},
                     {V, tri(dir, s, {1, 5, 4, 5, 17, 14})}, {V, tri(dir, s, {1, 5, 4, 5, 8, 17})}}, 0);
        add_term(P, {{V, tri(dir, s, {1, 2, 5, 5, 15, 14})}, {V, tri(dir, s, {1, 2, 5, 5, 6, 15})},
                     {V, tri(dir, s, {1, 5, 4, 5, 15, 14})}, {V, tri(dir, s, {1, 5, 4, 5, 6, 15})}}, 0);
    }
    const int n = ord;
    std::vector<double> gx, gw;
    gl_pm1(n, gx, gw);
    const int64_t ns = 24LL * n * n, ne = 6LL * n * n, nv = (int64_t)n * n * n;
    const int cs = (int)P.s.size(), ce = (int)P.e.size(), cv = (int)P.v.size();
    DevMem dg, dts, dte, dtv, dsng, dps, dpe, dpv, dout;
    SNG_HIP(hipMalloc(&dg.p, sizeof(double) * 2 * n));
    SNG_HIP(hipMemcpy(dg.p, gx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    SNG_HIP(hipMemcpy((double*)dg.p + n, gw.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    SNG_HIP(hipMalloc(&dts.p, sizeof(Tri2) * cs));
    SNG_HIP(hipMalloc(&dte.p, sizeof(Tri2) * ce));
    SNG_HIP(hipMalloc(&dtv.p, sizeof(Tri2) * cv));
    SNG_HIP(hipMalloc(&dsng.p, sizeof(int) * cv));
    SNG_HIP(hipMemcpy(dts.p, P.s.data(), sizeof(Tri2) * cs, hipMemcpyHostToDevice));
    SNG_HIP(hipMemcpy(dte.p, P.e.data(), sizeof(Tri2) * ce, hipMemcpyHostToDevice));
    SNG_HIP(hipMemcpy(dtv.p, P.v.data(), sizeof(Tri2) * cv, hipMemcpyHostToDevice));
    SNG_HIP(hipMemcpy(dsng.p, P.vsng.data(), sizeof(int) * cv, hipMemcpyHostToDevice));
    SNG_HIP(hipMalloc(&dps.p, sizeof(double2) * ns * cs));
    SNG_HIP(hipMalloc(&dpe.p, sizeof(double2) * ne * ce));
    SNG_HIP(hipMalloc(&dpv.p, sizeof(double2) * nv * cv));
    SNG_HIP(hipMalloc(&dout.p, sizeof(double2) * (cs + ce + cv)));
    const cd zeta(zr, zi);
    const cd c4 = 1.0 / (4.0 * M_PI * zeta * zeta);
    const Zeta z{2 * M_PI * zr, 2 * M_PI * zi, c4.real(), c4.imag()};
    const double* x = (const double*)dg.p;
    const double* w = x + n;
    const int thr = 64;
    hipLaunchKernelGGL(k_dfn_s, dim3((unsigned)((ns + thr - 1) / thr), (unsigned)cs), dim3(thr), 0, nullptr,
                       (const Tri2*)dts.p, n, x, w, z, zr, zi, (double2*)dps.p);
    SNG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_dfn_e, dim3((unsigned)((ne + thr - 1) / thr), (unsigned)ce), dim3(thr), 0, nullptr,
                       (const Tri2*)dte.p, n, x, w, z, zr, zi, (double2*)dpe.p);
    SNG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_dfn_v, dim3((unsigned)((nv + thr - 1) / thr), (unsigned)cv), dim3(thr), 0, nullptr,
                       (const Tri2*)dtv.p, (const int*)dsng.p, n, x, w, z, zr, zi, (double2*)dpv.p);
    SNG_HIP(hipGetLastError());
    double2* out = (double2*)dout.p;
    hipLaunchKernelGGL(k_dfn_sum, dim3((unsigned)cs), dim3(256), 0, nullptr, (const double2*)dps.p, ns, out);
    hipLaunchKernelGGL(k_dfn_sum, dim3((unsigned)ce), dim3(256), 0, nullptr, (const double2*)dpe.p, ne, out + cs);
    hipLaunchKernelGGL(k_dfn_sum, dim3((unsigned)cv), dim3(256), 0, nullptr, (const double2*)dpv.p, nv,
                       out + cs + ce);
    SNG_HIP(hipGetLastError());
    std::vector<double2> h((size_t)(cs + ce + cv));
    SNG_HIP(hipMemcpy(h.data(), out, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
#undef SNG_HIP
    // instance totals with their Jacobians (eqvJacS / eqvJacEV) and the
    // vertex reduction's pi^2 / 144
    auto crossn = [](const double a[3], const double b[3], const double c[3]) {
        double u[3], v[3];
        for (int d = 0; d < 3; ++d) { u[d] = b[d] - a[d]; v[d] = c[d] - a[d]; }
        const double X = u[1] * v[2] - u[2] * v[1], Y = u[2] * v[0] - u[0] * v[2], Zz = u[0] * v[1] - u[1] * v[0];
        return std::sqrt(X * X + Y * Y + Zz * Zz);
    };
    auto val = [&](int kind, int i) -> cd {
        if (kind == 0) {
            const Tri2& T = P.s[(size_t)i];
            const double c = crossn(T.r[0], T.r[1], T.r[2]);
            return cd(h[(size_t)i].x, h[(size_t)i].y) * (c * c / 12.0);
        }
        const Tri2& T = kind == 1 ? P.e[(size_t)i] : P.v[(size_t)i];
        const double j = crossn(T.r[0], T.r[1], T.r[2]) * crossn(T.r[3], T.r[4], T.r[5]) / 12.0;
        const size_t off = kind == 1 ? (size_t)cs : (size_t)(cs + ce);
        const cd a(h[off + (size_t)i].x, h[off + (size_t)i].y);
        return kind == 1 ? a * j : a * (j * M_PI * M_PI / 144.0);
    };
    auto term = [&](int t) {
        cd acc = 0;
        for (auto& pr : P.terms[(size_t)t]) acc += val(pr.first, pr.second);
        return acc;
    };
    const double vol = s[0] * s[1] * s[2];
    // wS: [xx yy zz]
    w18[0] = (term(0) / 2.0 + r_srf_slf(s[1], s[2], zeta)) / vol;
    w18[1] = (term(1) / 2.0 + r_srf_slf(s[0], s[2], zeta)) / vol;
    w18[2] = (term(2) / 2.0 + r_srf_slf(s[0], s[1], zeta)) / vol;
    // wE: [xxY xxZ yyX yyZ zzX zzY xy xz yz]; terms 3+4*j+k for dir index j
    const cd xxY = term(3) + r_srf_edg_flt(s[2], s[1], zeta), xxZ = term(5) + r_srf_edg_flt(s[1], s[2], zeta);
    const cd xyA = term(4) + r_srf_edg_crn(s[2], s[1], s[0], zeta), xzA = term(6) + r_srf_edg_crn(s[1], s[2], s[0], zeta);
    const cd yyZ = term(7) + r_srf_edg_flt(s[0], s[2], zeta), yyX = term(9) + r_srf_edg_flt(s[2], s[0], zeta);
    const cd yzA = term(8) + r_srf_edg_crn(s[0], s[2], s[1], zeta), xyB = term(10) + r_srf_edg_crn(s[2], s[1], s[0], zeta);
    const cd zzX = term(11) + r_srf_edg_flt(s[1], s[0], zeta), zzY = term(13) + r_srf_edg_flt(s[0], s[1], zeta);
    const cd xzB = term(12) + r_srf_edg_crn(s[1], s[2], s[0], zeta), yzB = term(14) + r_srf_edg_crn(s[0], s[2], s[1], zeta);
    const cd we[9] = {xxY, xxZ, yyX, yyZ, zzX, zzY, (xyA + xyB) / 2.0, (xzA + xzB) / 2.0, (yzA + yzB) / 2.0};
    for (int k = 0; k < 9; ++k) w18[3 + k] = we[k] / vol;
    // wV: [xx yy zz xy xz yz]; terms 15+3*j+k
    const cd xxO = term(15), xyVA = term(16), xzVA = term(17);
    const cd yyO = term(18), yzVA = term(19), xyVB = term(20);
    const cd zzO = term(21), xzVB = term(22), yzVB = term(23);
    const cd wv[6] = {xxO, yyO, zzO, (xyVA + xyVB) / 2.0, (xzVA + xzVB) / 2.0, (yzVA + yzVB) / 2.0};
    for (int k = 0; k < 6; ++k) w18[12 + k] = wv[k] / vol;
    for (int k = 0; k < 18; ++k)
        if (!std::isfinite(w18[k].real()) || !std::isfinite(w18[k].imag())) {
            err = "singular face-pair integrals contain non-numeric values";
            return GLA_ERR_NONFINITE;
        }
    return GLA_OK;
}

}  // namespace gla
