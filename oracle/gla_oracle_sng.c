/*
 * TEST INFRASTRUCTURE ONLY (part of the CPU oracle; nothing in the product
 * library links or calls it).
 *
 * Restatement of the reference's singular face-pair integrals for touching
 * cells, in FP64:
 *   - sclEgoN, the scalar Green function with its static part removed
 *     (src/glaInt.jl:51-59);
 *   - the DIRECTFN_E reductions of the weakly singular triangle-pair
 *     integrals: wekSInt / wekEInt / wekVInt and their helpers ψlim*, ηlim*,
 *     nS, nE, intNE, aS, aE, subTri, spx!, spxE!, spxV!, eqvJac*
 *     (src/glaIntSup.jl:13-469);
 *   - the square-panel heads wekS / wekE / wekV with wekSDir / wekEDir /
 *     wekVDir and the point grid wekGrdPts! (src/glaInt.jl:75-267, 356-398);
 *   - the closed-form static panel integrals rSrfSlf / rSrfEdgCrn /
 *     rSrfEdgFlt (src/glaInt.jl:410-488);
 *   - the Gauss-Legendre rule of gauQud (src/glaGen.jl:685), from the
 *     textbook Newton iteration on the Legendre polynomials (FastGaussQuadrature
 *     1.0.2 in the reference; same nodes and weights to rounding).
 * The triangle vertex orders passed to the reductions are the reference's,
 * including those whose shared vertex or edge is not the one the reduction's
 * polar transform is centred on: the kernels there are the regularised
 * (bounded) one except in wekVDir, and the values are what the reference
 * computes at a given order.
 *
 * Independent pins (tests/test_oracle_singular.py): the static closed forms
 * against the textbook coincident-square value and against a Duffy-type
 * integration over difference coordinates of the 1/R kernel; every wS / wE /
 * wV value against the same independent integration of the full kernel
 * (convergence with the order); the order-32 against order-40 agreement of
 * test/intConTest.jl.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double _Complex zc;

static const double SQ3 = 1.7320508075688772935274463415059;

/* GL rule on [-1,1]: x ascending-free order as produced, w weights */
static void gl_pm1(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        /* ascending order, as gausslegendre returns them */
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1 - z * z) * pp * pp);
    }
}

int gor_gl_pm1(int32_t n, double *x, double *w) { gl_pm1(n, x, w); return 0; }

/* ------------------------------------------------------------------------ */
/* kernels                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    const double *x, *w;
    zc zeta;
} Q;

/* sclEgo (src/glaInt.jl:41-44) */
static inline zc scl_ego(double d, zc z)
{
    return cexp(2.0 * M_PI * I * d * z) / (4.0 * M_PI * d * z * z);
}

/* sclEgoN (src/glaInt.jl:51-59) */
static inline zc scl_ego_n(double d, zc z)
{
    if (d > 1e-7) return (cexp(2.0 * M_PI * I * d * z) - 1.0) / (4.0 * M_PI * d * z * z);
    return ((I / z) - M_PI * d) / 2.0;
}

/* θf: the affine map of a GL node from [-1,1] to [a,b] */
static inline double thf(double a, double b, double p) { return 0.5 * ((b - a) * p + a + b); }

/* spx!: barycentric coordinates of two points of the standard simplex */
static inline void spx(double X[2][3], double e1, double e2, double x1, double x2)
{
    X[0][0] = (SQ3 * (1.0 - e1) - x1) / (2.0 * SQ3);
    X[0][1] = (SQ3 * (1.0 + e1) - x1) / (2.0 * SQ3);
    X[0][2] = x1 / SQ3;
    X[1][0] = (SQ3 * (1.0 - e2) - x2) / (2.0 * SQ3);
    X[1][1] = (SQ3 * (1.0 + e2) - x2) / (2.0 * SQ3);
    X[1][2] = x2 / SQ3;
}

/* distance between barycentric point X[0] of triangle (r0,r1,r2) and X[1] of
 * triangle (r3,r4,r5); same triangle when `self` (kerSN) */
static inline double bary_dist(const double r[6][3], const double X[2][3], int self)
{
    const int o = self ? 0 : 3;
    double s = 0;
    for (int d = 0; d < 3; ++d) {
        double v = X[0][0] * r[0][d] + X[0][1] * r[1][d] + X[0][2] * r[2][d] -
                   (X[1][0] * r[o][d] + X[1][1] * r[o + 1][d] + X[1][2] * r[o + 2][d]);
        s += v * v;
    }
    return sqrt(s);
}

static double cross_norm(const double a[3], const double b[3], const double c[3])
{
    double u[3], v[3];
    for (int d = 0; d < 3; ++d) { u[d] = b[d] - a[d]; v[d] = c[d] - a[d]; }
    double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    return sqrt(x * x + y * y + z * z);
}

/* eqvJacS (src/glaIntSup.jl:386-390): |(r1-r2) x (r3-r1)|^2 / 12 */
static double jac_s(const double r[6][3])
{
    double c = cross_norm(r[0], r[1], r[2]);
    return c * c / 12.0;
}

/* eqvJacEV (src/glaIntSup.jl:377-384) */
static double jac_ev(const double r[6][3]) { return cross_norm(r[0], r[1], r[2]) * cross_norm(r[3], r[4], r[5]) / 12.0; }

/* ------------------------------------------------------------------------ */
/* self triangle (wekSInt, src/glaIntSup.jl:13-39)                          */
/* ------------------------------------------------------------------------ */
static void psi_lim_s(int idf, double *a, double *b)
{
    if (idf == 1 || idf == 5 || idf == 6) { *a = 0.0; *b = M_PI / 3.0; }
    else if (idf == 2 || idf == 7) { *a = M_PI / 3.0; *b = 2.0 * M_PI / 3.0; }
    else { *a = 2.0 * M_PI / 3.0; *b = M_PI; }
}

static void eta_lim_s(int idf, double th, double *a, double *b)
{
    switch (idf) {
    case 1: case 2: *a = 0.0; *b = 1.0; break;
    case 3: *a = (SQ3 - tan(M_PI - th)) / (SQ3 + tan(M_PI - th)); *b = 1.0; break;
    case 4: *a = 0.0; *b = (SQ3 - tan(M_PI - th)) / (SQ3 + tan(M_PI - th)); break;
    case 5: *a = (tan(th) - SQ3) / (SQ3 + tan(th)); *b = 0.0; break;
    case 6: *a = -1.0; *b = (tan(th) - SQ3) / (SQ3 + tan(th)); break;
    default: *a = -1.0; *b = 0.0; break;   /* 7, 8 */
    }
}

/* subTri (src/glaIntSup.jl:364-375) */
static inline void sub_tri(double l1, double l2, int dir, double *e, double *x)
{
    if (dir == 1) { *e = l1; *x = l2; }
    else if (dir == 2) { *e = 0.5 * (1.0 - l1 - l2 * SQ3); *x = 0.5 * (SQ3 + l1 * SQ3 - l2); }
    else { *e = 0.5 * (-1.0 - l1 + l2 * SQ3); *x = 0.5 * (SQ3 - l1 * SQ3 - l2); }
}

/* aS (src/glaIntSup.jl:320-340) */
static zc a_s(const double r[6][3], double t1, double tb, double th, int dir, const Q *q)
{
    const double s1 = sin(t1), c1 = cos(t1);
    double e1, x1;
    sub_tri(tb, th * s1, dir, &e1, &x1);
    zc acc = 0;
    for (int n = 0; n < q->n; ++n) {
        const double t = thf(0.0, th, q->x[n]);
        double e2, x2;
        sub_tri(t * c1 + tb, (th - t) * s1, dir, &e2, &x2);
        double X[2][3];
        spx(X, e1, e2, x1, x2);
        acc += q->w[n] * t * scl_ego_n(bary_dist(r, X, 1), q->zeta);
    }
    return 0.5 * th * acc;
}

/* nS (src/glaIntSup.jl:200-237) */
static zc n_s(const double r[6][3], int dir, int idf, double t1, double tb, const Q *q)
{
    double lim, pre;
    if (idf == 1 || idf == 5) { lim = (1.0 - tb) / cos(t1); pre = (1.0 - tb) / (2.0 * cos(t1)); }
    else if (idf == 2 || idf == 3) { lim = SQ3 * (1.0 - tb) / sin(t1); pre = SQ3 * (1.0 - tb) / (2.0 * sin(t1)); }
    else if (idf == 6 || idf == 7) { lim = SQ3 * (1.0 + tb) / sin(t1); pre = SQ3 * (1.0 + tb) / (2.0 * sin(t1)); }
    else { lim = -(1.0 + tb) / cos(t1); pre = -(1.0 + tb) / (2.0 * cos(t1)); }
    zc acc = 0;
    for (int n = 0; n < q->n; ++n) acc += q->w[n] * a_s(r, t1, tb, thf(0.0, lim, q->x[n]), dir, q);
    return pre * acc;
}

/* sum of per-item partials in item order (the items run in parallel) */
static zc sum_items(const zc *part, int n)
{
    zc acc = 0;
    for (int i = 0; i < n; ++i) acc += part[i];
    return acc;
}

static zc wek_s_int(const double r[6][3], const Q *q)
{
    /* CartesianIndices((1:3, 1:8, 1:n)): first index fastest */
    const int ni = 24 * q->n;
    zc *part = (zc *)malloc(sizeof(zc) * (size_t)ni);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < ni; ++i) {
        const int dir = i % 3 + 1, idf = (i / 3) % 8 + 1, k = i / 24;
        double pa, pb, ea, eb;
        psi_lim_s(idf, &pa, &pb);
        const double th = thf(pa, pb, q->x[k]);
        eta_lim_s(idf, th, &ea, &eb);
        zc v = 0;
        for (int it = 0; it < q->n; ++it) v += q->w[it] * n_s(r, dir, idf, th, thf(ea, eb, q->x[it]), q);
        part[i] = 0.25 * v * q->w[k] * (pb - pa) * (eb - ea) * sin(th);
    }
    const zc acc = sum_items(part, ni);
    free(part);
    return jac_s(r) * acc;
}

/* ------------------------------------------------------------------------ */
/* edge-adjacent triangles (wekEInt, src/glaIntSup.jl:43-68)                */
/* ------------------------------------------------------------------------ */
static void psi_lim_e(int idf, double *a, double *b)
{
    if (idf == 1) { *a = 0.0; *b = M_PI / 3.0; }
    else if (idf == 2 || idf == 3) { *a = M_PI / 3.0; *b = M_PI / 2.0; }
    else if (idf == 4 || idf == 6) { *a = M_PI / 2.0; *b = M_PI; }
    else { *a = 0.0; *b = M_PI / 2.0; }   /* 5 */
}

static void eta_lim_e(int idf, double th, double *a, double *b)
{
    const double s = sin(th), c = cos(th);
    switch (idf) {
    case 1: *a = 0.0; *b = atan(s + SQ3 * c); break;
    case 2: *a = atan(s - SQ3 * c); *b = atan(s + SQ3 * c); break;
    case 3: case 4: *a = 0.0; *b = atan(s - SQ3 * c); break;
    case 5: *a = atan(s + SQ3 * c); *b = 0.5 * M_PI; break;
    default: *a = atan(s - SQ3 * c); *b = 0.5 * M_PI; break;   /* 6 */
    }
}

/* spxE! (src/glaIntSup.jl:409-425) */
static inline void spx_e(double X[2][3], double lam, double eta, double tb, double t1, int idf)
{
    const double sb = sin(tb), cb = cos(tb), s1 = sin(t1), c1 = cos(t1);
    if (idf == 1) spx(X, eta, lam * cb * c1 - eta, lam * sb, lam * cb * s1);
    else spx(X, -eta, -(lam * cb * c1 - eta), lam * sb, lam * cb * s1);
}

/* aE (src/glaIntSup.jl:342-357) */
static zc a_e(const double r[6][3], double lam, double eta, double tb, double t1, int idf, const Q *q)
{
    zc acc = 0;
    for (int n = 0; n < q->n; ++n) {
        const double z = thf(0.0, lam, q->x[n]);
        double X[2][3];
        spx_e(X, z, eta, tb, t1, idf);
        acc += q->w[n] * z * z * scl_ego_n(bary_dist(r, X, 0), q->zeta);
    }
    return 0.5 * lam * acc;
}

/* intNE (src/glaIntSup.jl:289-318) */
static zc int_ne(const double r[6][3], int n, int k, double g, double tb, double t1, int idf2, const Q *q)
{
    double eta, lam;
    const double s1 = sin(t1), c1 = cos(t1);
    switch (k) {
    case 1: eta = thf(0.0, g, q->x[n]); lam = SQ3 * (1 + eta) / (cos(tb) * (s1 + SQ3 * c1)); break;
    case 2: eta = thf(g, 1.0, q->x[n]); lam = SQ3 * (1.0 - fabs(eta)) / sin(tb); break;
    case 3: eta = thf(g, 1.0, q->x[n]); lam = SQ3 * (1.0 - fabs(eta)) / (cos(tb) * (s1 - SQ3 * c1)); break;
    case 4: eta = thf(0.0, 1.0, q->x[n]); lam = SQ3 * (1.0 - fabs(eta)) / (cos(tb) * (s1 - SQ3 * c1)); break;
    default: eta = thf(0.0, 1.0, q->x[n]); lam = SQ3 * (1.0 - eta) / sin(tb); break;   /* 5 */
    }
    return a_e(r, lam, eta, tb, t1, idf2, q);
}

/* nE (src/glaIntSup.jl:239-287): tb and t1 are its third and fourth
 * arguments (named θB and θ1 there) */
static zc n_e(const double r[6][3], int idf1, int idf2, double tb, double t1, const Q *q)
{
    zc i1 = 0, i2 = 0;
    if (idf1 == 1 || idf1 == 2) {
        const double s1 = sin(t1), c1 = cos(t1);
        const double g = (s1 + SQ3 * c1 - tan(tb)) / (s1 + SQ3 * c1 + tan(tb));
        for (int n = 0; n < q->n; ++n) {
            i1 += q->w[n] * int_ne(r, n, 1, g, tb, t1, idf2, q);
            i2 += q->w[n] * int_ne(r, n, 2, g, tb, t1, idf2, q);
        }
        return 0.5 * i2 + g * 0.5 * (i1 - i2);
    } else if (idf1 == 3) {
        const double g = SQ3 / tan(t1);
        for (int n = 0; n < q->n; ++n) {
            i1 += q->w[n] * int_ne(r, n, 1, g, tb, t1, idf2, q);
            i2 += q->w[n] * int_ne(r, n, 3, g, tb, t1, idf2, q);
        }
        return 0.5 * i2 + 0.5 * g * (i1 - i2);
    } else if (idf1 == 4) {
        for (int n = 0; n < q->n; ++n) i1 += q->w[n] * int_ne(r, n, 4, 1.0, tb, t1, idf2, q);
        return 0.5 * i1;
    }
    for (int n = 0; n < q->n; ++n) i1 += q->w[n] * int_ne(r, n, 5, 1.0, tb, t1, idf2, q);
    return 0.5 * i1;
}

static zc wek_e_int(const double r[6][3], const Q *q)
{
    /* CartesianIndices((1:6, 1:n)) */
    const int ni = 6 * q->n;
    zc *part = (zc *)malloc(sizeof(zc) * (size_t)ni);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < ni; ++i) {
        const int idf = i % 6 + 1, k = i / 6;
        double pa, pb, ea, eb;
        psi_lim_e(idf, &pa, &pb);
        const double thb = thf(pa, pb, q->x[k]);
        eta_lim_e(idf, thb, &ea, &eb);
        zc v = 0;
        for (int it = 0; it < q->n; ++it) {
            const double tha = thf(ea, eb, q->x[it]);
            v += q->w[it] * cos(tha) * (n_e(r, idf, 1, tha, thb, q) + n_e(r, idf, -1, tha, thb, q));
        }
        part[i] = 0.25 * v * q->w[k] * (pb - pa) * (eb - ea);
    }
    const zc acc = sum_items(part, ni);
    free(part);
    return jac_ev(r) * acc;
}

/* ------------------------------------------------------------------------ */
/* vertex-adjacent triangles (wekVInt, src/glaIntSup.jl:72-135)             */
/* ------------------------------------------------------------------------ */
/* spxV! (src/glaIntSup.jl:398-407) */
static inline void spx_v(double X[2][3], double t4, double t3, double tb, double t1)
{
    const double sb = sin(tb), cb = cos(tb), s1 = sin(t1), c1 = cos(t1), s3 = sin(t3), c3 = cos(t3);
    spx(X, t4 * c3 * c1 - 1.0, t4 * s3 * cb - 1.0, t4 * c3 * s1, t4 * s3 * sb);
}

static zc wek_v_int(int sng, const double r[6][3], const Q *q)
{
    /* CartesianIndices((1:n, 1:n, 1:n)); one partial per (i2, i3) */
    const int ni = q->n * q->n;
    zc *part = (zc *)malloc(sizeof(zc) * (size_t)ni);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ni; ++j) {
        const int i2 = j % q->n, i3 = j / q->n;
        zc acc = 0;
            for (int i1 = 0; i1 < q->n; ++i1) {
                const double ta = thf(0.0, M_PI / 3.0, q->x[i1]);
                const double la = 2.0 * SQ3 / (sin(ta) + SQ3 * cos(ta));
                const double tb = thf(0.0, M_PI / 3.0, q->x[i2]);
                const double lb = 2.0 * SQ3 / (sin(tb) + SQ3 * cos(tb));
                const double at = atan(lb / la);
                const double tc = thf(0.0, at, q->x[i3]);
                const double td = thf(at, M_PI / 2.0, q->x[i3]);
                const double sc = sin(tc), cc = cos(tc), sd = sin(td), cd = cos(td);
                zc ic = 0, id = 0;
                for (int c = 0; c < q->n; ++c) {
                    const double tx = thf(0.0, la / cc, q->x[c]);
                    double X[2][3];
                    spx_v(X, tx, tc, tb, ta);
                    const double d = bary_dist(r, X, 0);
                    ic += q->w[c] * tx * tx * tx * (sng ? scl_ego_n(d, q->zeta) : scl_ego(d, q->zeta));
                }
                for (int e = 0; e < q->n; ++e) {
                    const double tx = thf(0.0, lb / sd, q->x[e]);
                    double X[2][3];
                    spx_v(X, tx, td, tb, ta);
                    const double d = bary_dist(r, X, 0);
                    id += q->w[e] * tx * tx * tx * (sng ? scl_ego_n(d, q->zeta) : scl_ego(d, q->zeta));
                }
                acc += q->w[i1] * q->w[i2] * q->w[i3] * (at * (la * sc * ic - lb * cd * id) + 0.5 * M_PI * lb * cd * id);
            }
        part[j] = acc;
    }
    const zc acc = sum_items(part, ni);
    free(part);
    return jac_ev(r) * M_PI * M_PI * acc / 144.0;
}

/* ------------------------------------------------------------------------ */
/* square panels                                                            */
/* ------------------------------------------------------------------------ */
/* wekGrdPts! (src/glaInt.jl:356-398): P[k] is the reference's grdPts[:,k+1] */
static void grd_pts(int dir, const double s[3], double P[18][3])
{
    double gx, gy, gz;
    if (dir == 1) { gx = s[0]; gy = s[1]; gz = s[2]; }
    else if (dir == 2) { gx = s[2]; gy = s[0]; gz = s[1]; }
    else { gx = s[1]; gy = s[2]; gz = s[0]; }
    for (int k = 0; k < 18; ++k) {
        P[k][0] = (k % 3) * gx;
        P[k][1] = ((k / 3) % 3) * gy;
        P[k][2] = (k / 9) * gz;
    }
}

/* triangle-pair argument from 1-based point numbers of the grid */
static void tri2(const double P[18][3], int a, int b, int c, int d, int e, int f, double r[6][3])
{
    const int id[6] = {a, b, c, d, e, f};
    for (int k = 0; k < 6; ++k) memcpy(r[k], P[id[k] - 1], sizeof(double) * 3);
}

static zc S3(const double P[18][3], int a, int b, int c, const Q *q)
{
    double r[6][3];
    tri2(P, a, b, c, a, b, c, r);
    return wek_s_int(r, q);
}

static zc E6(const double P[18][3], int a, int b, int c, int d, int e, int f, const Q *q)
{
    double r[6][3];
    tri2(P, a, b, c, d, e, f, r);
    return wek_e_int(r, q);
}

static zc V6(int sng, const double P[18][3], int a, int b, int c, int d, int e, int f, const Q *q)
{
    double r[6][3];
    tri2(P, a, b, c, d, e, f, r);
    return wek_v_int(sng, r, q);
}

/* wekSDir (src/glaInt.jl:92-116) */
static zc wek_s_dir(int dir, const double s[3], const Q *q)
{
    double P[18][3];
    grd_pts(dir, s, P);
    return ((S3(P, 1, 2, 5, q) + S3(P, 1, 5, 4, q) + E6(P, 1, 2, 5, 1, 5, 4, q) + E6(P, 1, 5, 4, 1, 2, 5, q)) +
            (S3(P, 4, 1, 2, q) + S3(P, 4, 2, 5, q) + E6(P, 4, 1, 2, 4, 2, 5, q) + E6(P, 4, 2, 5, 4, 1, 2, q))) /
           2.0;
}

/* wekEDir (src/glaInt.jl:166-199) */
static void wek_e_dir(int dir, const double s[3], const Q *q, zc v[4])
{
    double P[18][3];
    grd_pts(dir, s, P);
    v[0] = E6(P, 1, 2, 5, 2, 3, 5, q) + V6(1, P, 1, 2, 5, 3, 6, 5, q) + V6(1, P, 1, 5, 4, 2, 3, 5, q) +
           V6(1, P, 1, 5, 4, 3, 6, 5, q);
    v[1] = E6(P, 1, 2, 5, 2, 11, 5, q) + V6(1, P, 1, 2, 5, 11, 14, 5, q) + V6(1, P, 1, 5, 4, 2, 11, 5, q) +
           V6(1, P, 1, 5, 4, 11, 14, 5, q);
    v[2] = V6(1, P, 1, 2, 5, 4, 5, 7, q) + V6(1, P, 1, 2, 5, 5, 8, 7, q) + E6(P, 1, 5, 4, 4, 5, 7, q) +
           V6(1, P, 1, 5, 4, 5, 8, 7, q);
    v[3] = V6(1, P, 1, 2, 5, 4, 5, 13, q) + V6(1, P, 1, 2, 5, 5, 14, 13, q) + E6(P, 1, 5, 4, 4, 5, 13, q) +
           V6(1, P, 1, 5, 4, 5, 14, 13, q);
}

/* wekVDir (src/glaInt.jl:233-259) */
static void wek_v_dir(int dir, const double s[3], const Q *q, zc v[3])
{
    double P[18][3];
    grd_pts(dir, s, P);
    v[0] = V6(0, P, 1, 2, 5, 5, 6, 9, q) + V6(0, P, 1, 2, 5, 5, 9, 8, q) + V6(0, P, 1, 5, 4, 5, 6, 9, q) +
           V6(0, P, 1, 5, 4, 5, 9, 8, q);
    v[1] = V6(0, P, 1, 2, 5, 5, 17, 14, q) + V6(0, P, 1, 2, 5, 5, 8, 17, q) + V6(0, P, 1, 5, 4, 5, 17, 14, q) +
           V6(0, P, 1, 5, 4, 5, 8, 17, q);
    v[2] = V6(0, P, 1, 2, 5, 5, 15, 14, q) + V6(0, P, 1, 2, 5, 5, 6, 15, q) + V6(0, P, 1, 5, 4, 5, 15, 14, q) +
           V6(0, P, 1, 5, 4, 5, 6, 15, q);
}

/* ------------------------------------------------------------------------ */
/* closed-form static panel integrals (src/glaInt.jl:410-488)               */
/* ------------------------------------------------------------------------ */
/* rSrfSlf: 1/(4 pi R) over a la x lb rectangle with itself */
zc gor_r_srf_slf(double la, double lb, zc z)
{
    const double a2 = la * la, b2 = lb * lb;
    const double r1 = sqrt(a2 + b2), r4 = sqrt(a2 + 4 * b2);
    const double t1 = 8 * la * a2 + 8 * lb * b2 - 8 * a2 * r1 - 8 * b2 * r1;
    const double t2 = -3 * a2 * lb *
                      (2 * log(la) + 2 * log(la + lb - r1) + log(r1 - lb) - 5 * log(lb + r1) - 2 * log(lb - la + r1) -
                       2 * log(la + 2 * lb - r4) + log(r4 - 2 * lb) + 2 * log(la - 2 * lb + r4) + log(2 * lb + r4) +
                       2 * log(2 * lb - la + r4) - 2 * log(la + 2 * lb + r4));
    const double t3 = 6 * la * b2 *
                      (log(64.0) + 4 * log(lb) + 2 * log(r1 - la) + 3 * log(la + r1) - 3 * log(r4 - la) -
                       3 * log(sqrt(a2 * a2 + 5 * a2 * b2 + 4 * b2 * b2) + la * (r1 - la - r4)));
    return (1.0 / (48.0 * M_PI * z * z)) * (t1 + t2 + t3);
}

/* rSrfEdgCrn: two rectangles at a right angle sharing an edge of length la;
 * lb and lc their other sides */
zc gor_r_srf_edg_crn(double la, double lb, double lc, zc z)
{
    const double a2 = la * la, b2 = lb * lb, c2 = lc * lc, a3 = a2 * la, c3 = c2 * lc;
    const double R = sqrt(a2 + b2 + c2), rbc = sqrt(b2 + c2);
    double v = 8 * lb * lc * rbc - 8 * lb * lc * R - 12 * a3 * atan(1.0 / (la * lc / (a2 + b2 - lb * R))) +
               12 * a3 * atan(la / lc) - 12 * la * c2 * atan(la * lb / (lc * R)) -
               12 * la * b2 * atan(la * lc / (lb * R)) - 16 * a3 * atan(lb * lc / (la * R)) +
               6 * c3 * atanh(lb / rbc) - 6 * lc * (a2 + c2) * atanh(lb / R) - 15 * a2 * lc * log(a2 + c2) -
               c3 * log(a2 + c2) + 2 * c3 * log(lc / (lb + rbc)) + 6 * a2 * lc * log(R - lb) +
               24 * a2 * lc * log(R + lb) + 2 * c3 * log(R + lb) +
               6 * la * lb *
                   (-2 * la * log(a2 + b2) - lc * log((b2 + c2) * (R - la)) + 3 * lc * log(la + R) +
                    la * log(R - lc) + 3 * la * log(R + lc)) +
               2 * lb * b2 * (log((R - lc) / (lc + R)) + log(1 + (2 * lc * (lc + rbc)) / b2));
    return (1.0 / (48.0 * M_PI * z * z)) * v;
}

/* rSrfEdgFlt: two coplanar la x lb rectangles sharing an edge of length la.
 * As in the reference, the logarithmic part carries 1/(64 pi) without the
 * 1/zeta^2 of the other terms (the two agree at zeta = 1). */
zc gor_r_srf_edg_flt(double la, double lb, zc z)
{
    const double a2 = la * la, b2 = lb * lb;
    const double r1 = sqrt(a2 + b2), r4 = sqrt(a2 + 4 * b2);
    const zc p1 = (1.0 / (12.0 * M_PI * z * z)) * (-la * a2 + 2 * b2 * (3 * lb + r1 - 2 * r4) + a2 * (2 * r1 - r4));
    const double q1 = lb * (-62 * log(2.0) - 5 * log(-la + r1) + 4 * log(8 * b2 * (-la + r1)) - 33 * log(la + r1) +
                            17 * log(-la + r4) - 24 * log(lb * (-la + r4)) + 57 * log(la + r4));
    const double q2 =
        4 * la *
        (-8 * asinh(lb / la) + 6 * asinh(2 * lb / la) + 6 * atanh(lb / r1) + 12 * log(la) - 13 * log(-lb + r1) +
         log((-lb + r1) / la) + log(la / (lb + r1)) - 7 * log(lb + r1) - 2 * log((lb + r1) / la) -
         3 * log(-(((lb + r1) * (2 * lb - r4)) / a2)) - 3 * log((-lb + r1) / (-2 * lb + r4)) + 11 * log(-2 * lb + r4) -
         3 * log((lb + r1) / (2 * lb + r4)) + log(2 * lb + r4) + 9 * log((2 * lb + r4) / (lb + r1)) -
         2 * log(a2 + 2 * lb * (lb - r1)));
    return p1 + (1.0 / (64.0 * M_PI)) * la * lb * (q1 + q2);
}

/* ------------------------------------------------------------------------ */
/* heads: wekS / wekE / wekV (src/glaInt.jl:75-89, 123-155, 204-226)        */
/* ------------------------------------------------------------------------ */
static void make_q(int ord, zc zeta, Q *q, double **buf)
{
    double *b = (double *)malloc(sizeof(double) * 2 * (size_t)ord);
    gl_pm1(ord, b, b + ord);
    q->n = ord;
    q->x = b;
    q->w = b + ord;
    q->zeta = zeta;
    *buf = b;
}

/* wS = [xx yy zz] (face normal), values not yet divided by the cell volume */
int gor_wek_s(const double s[3], double zr, double zi, int32_t ord, zc out[3])
{
    Q q;
    double *b;
    const zc z = zr + I * zi;
    make_q(ord, z, &q, &b);
    out[0] = wek_s_dir(3, s, &q) + gor_r_srf_slf(s[1], s[2], z);
    out[1] = wek_s_dir(2, s, &q) + gor_r_srf_slf(s[0], s[2], z);
    out[2] = wek_s_dir(1, s, &q) + gor_r_srf_slf(s[0], s[1], z);
    free(b);
    return 0;
}

/* wE = [xxY xxZ yyX yyZ zzX zzY xy xz yz] */
int gor_wek_e(const double s[3], double zr, double zi, int32_t ord, zc out[9])
{
    Q q;
    double *b;
    const zc z = zr + I * zi;
    make_q(ord, z, &q, &b);
    zc v[4];
    wek_e_dir(3, s, &q, v);
    const zc xxY = v[0] + gor_r_srf_edg_flt(s[2], s[1], z);
    const zc xxZ = v[2] + gor_r_srf_edg_flt(s[1], s[2], z);
    const zc xyA = v[1] + gor_r_srf_edg_crn(s[2], s[1], s[0], z);
    const zc xzA = v[3] + gor_r_srf_edg_crn(s[1], s[2], s[0], z);
    wek_e_dir(2, s, &q, v);
    const zc yyZ = v[0] + gor_r_srf_edg_flt(s[0], s[2], z);
    const zc yyX = v[2] + gor_r_srf_edg_flt(s[2], s[0], z);
    const zc yzA = v[1] + gor_r_srf_edg_crn(s[0], s[2], s[1], z);
    const zc xyB = v[3] + gor_r_srf_edg_crn(s[2], s[1], s[0], z);
    wek_e_dir(1, s, &q, v);
    const zc zzX = v[0] + gor_r_srf_edg_flt(s[1], s[0], z);
    const zc zzY = v[2] + gor_r_srf_edg_flt(s[0], s[1], z);
    const zc xzB = v[1] + gor_r_srf_edg_crn(s[1], s[2], s[0], z);
    const zc yzB = v[3] + gor_r_srf_edg_crn(s[0], s[2], s[1], z);
    out[0] = xxY; out[1] = xxZ; out[2] = yyX; out[3] = yyZ; out[4] = zzX; out[5] = zzY;
    out[6] = (xyA + xyB) / 2.0;
    out[7] = (xzA + xzB) / 2.0;
    out[8] = (yzA + yzB) / 2.0;
    free(b);
    return 0;
}

/* wV = [xx yy zz xy xz yz] */
int gor_wek_v(const double s[3], double zr, double zi, int32_t ord, zc out[6])
{
    Q q;
    double *b;
    make_q(ord, zr + I * zi, &q, &b);
    zc v[3];
    wek_v_dir(3, s, &q, v);
    const zc xxO = v[0], xyA = v[1], xzA = v[2];
    wek_v_dir(2, s, &q, v);
    const zc yyO = v[0], yzA = v[1], xyB = v[2];
    wek_v_dir(1, s, &q, v);
    const zc zzO = v[0], xzB = v[1], yzB = v[2];
    out[0] = xxO; out[1] = yyO; out[2] = zzO;
    out[3] = (xyA + xyB) / 2.0;
    out[4] = (xzA + xzB) / 2.0;
    out[5] = (yzA + yzB) / 2.0;
    free(b);
    return 0;
}

/* the static kernel values alone, for the closed-form pins */
int gor_r_srf(int32_t which, double la, double lb, double lc, double *out)
{
    zc v = which == 0 ? gor_r_srf_slf(la, lb, 1.0) : which == 1 ? gor_r_srf_edg_flt(la, lb, 1.0)
                                                                : gor_r_srf_edg_crn(la, lb, lc, 1.0);
    out[0] = creal(v);
    out[1] = cimag(v);
    return 0;
}
