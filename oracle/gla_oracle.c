/*
 * TEST INFRASTRUCTURE ONLY.  This is the CPU oracle that the tests, the smoke
 * check and the CPU-baseline leg of bench.py use as the checker.  Nothing in
 * the product library links, loads or calls it.
 *
 * Contents (citations are to the reference repository):
 *   - gor_matvec_c128 / gor_matvec_c64: restatement of the host matvec
 *     egoOpr! -> egoBrnHst! (src/glaAct.jl:33-183) with the split / merge /
 *     leaf-multiply helpers of src/glaActSup.jl:72-82, 163-217, 425-436,
 *     504-535, 671-766, 879-899 and the 1-D FFTs of src/glaAct.jl:88,170
 *     (FFTW there, a plain Stockham FFT here).
 *   - gor_fur_from_crc: restatement of the Fourier-storage construction in
 *     GlaOprMem (src/glaMemSup.jl:96-156): six 3-D FFTs of the circulant,
 *     even/odd branch extraction, symmetric truncation.
 *   - gor_crc_self: restatement of egoToeCrc! (src/glaGen.jl:143-163,
 *     indSel :647-654, indFlp :658-661).
 *   - gor_crc_ext: circulant over target-minus-source separations as in
 *     egoFunExt! / grdSel (src/glaGen.jl:169-189, 632-643).
 *   - gor_crc_prt / gor_tensor_mixed: the circulant of one (source, target)
 *     partition pair of an operator whose cell scales differ: genEgoExt! on
 *     the partition volumes of GlaOprMem (src/glaMemSup.jl:64-89), all 36
 *     face pairs of egoFunOut! (src/glaGen.jl:233-246) between cells of
 *     different sizes with the srfScl factors (src/glaInt.jl:284-305).
 *   - gor_brute_crc: independent direct circular sum y[t] = sum_s c[t-s] x[s]
 *     over the real-space circulant, used to pin the restatement.
 *   - gor_toe_model: test Green tensors (a point-collocation dyadic model,
 *     the face-pair surface-integral formulation with fixed-order rules, and
 *     model 3, the reference's construction: egoFunInn! for separated cells
 *     and egoFunSng! with the DIRECTFN singular values of gla_oracle_sng.c
 *     for touching ones; see the function headers).
 *
 * How each part is pinned (tests/test_oracle.py): the FFT against the DFT
 * definition and numpy; the matvec restatement against gor_brute_crc on self,
 * odd-size and external grids, and gor_brute_crc against a direct
 * block-Toeplitz sum built from the real-space tensor with the reflection
 * signs; the Fourier storage by the mirror identity of the stored branches;
 * linearity, reciprocity and the self-vs-external translation identity of
 * test/extSlfTest.jl:45-47; the face-pair tensor's far field against the
 * analytic dyadic of test/anaTest.jl:36-44 with the cube-averaging error
 * (kh)^2/12; model 3 (tests/test_oracle_singular.py) against an independent
 * difference-coordinate quadrature of every face pair: the closed-form static
 * panel integrals exactly, the singular values to their order-32 quadrature
 * error (<= 1e-10 self and edge, <= 2e-9 vertex), the tensor at the eight
 * touching offsets (the correction tables, anisotropic cells included) and at
 * regular offsets (<= 1e-11), plus the order-32 vs order-40 convergence of
 * test/intConTest.jl.  The mixed-size tensor (tests/test_oracle_partition.py)
 * is pinned by target and source subdivision: a cell of twice the size on
 * every axis is the sum of its 8 subcells (the face-pair sum is additive
 * over a subdivision, interior faces cancel, and srfScl reduces to 1 / V_T
 * for cells of one shape), so it must equal the corresponding sum of
 * equal-size model-3 tensors; with equal sizes it reduces to model 3
 * itself.  For a target cell twice as long on one axis only, srfScl's factors
 * differ per component (1/2 of the subcell sum on xx, yy, zz, yz, the whole
 * sum on xy, xz, from the way srfScl indexes its table), and the same
 * subdivision test checks that pattern.  The reference's hcubature values of the regular face
 * pairs (third-party, tolerance 1e-6, src/glaGen.jl:574-576) cannot be run
 * here: model 3 integrates those pairs to their limit instead, so its
 * agreement with the reference is bounded by that tolerance.
 *
 * Build: oracle/Makefile -> oracle/_build/libgila_oracle.so
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define R double
#define C double _Complex
#define SFX(x) x##_c128
#include "gla_oracle_impl.h"
#undef R
#undef C
#undef SFX
#undef IDX4

#define R float
#define C float _Complex
#define SFX(x) x##_c64
#include "gla_oracle_impl.h"
#undef R
#undef C
#undef SFX
#undef IDX4

typedef double _Complex zc;

/* ------------------------------------------------------------------------ */
/* dimensions                                                               */
/* ------------------------------------------------------------------------ */
/* B = (S+T)/2 (src/glaAct.jl:68); U as in src/glaMemSup.jl:121-134. */
int gor_dims(const int64_t S[3], const int64_t T[3], const int32_t sym[3], int64_t B[3], int64_t U[3])
{
    for (int d = 0; d < 3; ++d) {
        if (S[d] < 1 || T[d] < 1 || ((S[d] + T[d]) & 1)) return 1;
        B[d] = (S[d] + T[d]) / 2;
        if (sym[d]) {
            int64_t u = (T[d] + 1) / 2 + (T[d] % 2 == 0);
            U[d] = u < 2 ? 2 : u;
        } else {
            U[d] = B[d];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fourier storage from the real-space circulant                            */
/* ------------------------------------------------------------------------ */
/* crc: (3,3,L1,L2,L3) complex double, column-major.  fur: 8 consecutive
 * (U1,U2,U3,6) blocks.  Component order [xx,yy,zz,xy,xz,yz] from the upper
 * triangle (blkEgoItr, src/glaMemSup.jl:273-291). */
int gor_fur_from_crc(const int64_t S[3], const int64_t T[3], const int32_t sym[3],
                     const zc *crc, zc *fur)
{
    int64_t B[3], U[3];
    if (gor_dims(S, T, sym, B, U)) return 1;
    const int64_t L[3] = {2 * B[0], 2 * B[1], 2 * B[2]};
    const int64_t Lt = L[0] * L[1] * L[2];
    zc *prp = (zc *)malloc(sizeof(zc) * (size_t)Lt * 6);
    const int rows[6] = {0, 1, 2, 0, 0, 1}, cols[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; ++c) {
        zc *dst = prp + (int64_t)c * Lt;
        for (int64_t p = 0; p < Lt; ++p) dst[p] = crc[rows[c] + 3 * cols[c] + 9 * p];
        fft_axis_c128(dst, L, 1, 0, -1, 1.0);
        fft_axis_c128(dst, L, 1, 1, -1, 1.0);
        fft_axis_c128(dst, L, 1, 2, -1, 1.0);
    }
    for (int c = 0; c < 6; ++c)
        for (int64_t p = 0; p < Lt; ++p)
            if (isnan(creal(prp[c * Lt + p])) || isnan(cimag(prp[c * Lt + p])) ||
                isinf(creal(prp[c * Lt + p])) || isinf(cimag(prp[c * Lt + p]))) {
                free(prp);
                return 2;
            }
    const int64_t Ut = U[0] * U[1] * U[2];
    for (int eo = 0; eo < 8; ++eo) {
        const int par[3] = {(eo >> 2) & 1, (eo >> 1) & 1, eo & 1};
        zc *dst = fur + (int64_t)eo * Ut * 6;
        for (int c = 0; c < 6; ++c)
            for (int64_t k3 = 0; k3 < U[2]; ++k3)
                for (int64_t k2 = 0; k2 < U[1]; ++k2)
                    for (int64_t k1 = 0; k1 < U[0]; ++k1) {
                        /* a branch shorter than two entries is broadcast, as the
                         * reference's (max(B,2))-sized intermediate is */
                        int64_t q1 = k1 < B[0] ? k1 : B[0] - 1;
                        int64_t q2 = k2 < B[1] ? k2 : B[1] - 1;
                        int64_t q3 = k3 < B[2] ? k3 : B[2] - 1;
                        int64_t m1 = 2 * q1 + par[0], m2 = 2 * q2 + par[1], m3 = 2 * q3 + par[2];
                        dst[k1 + U[0] * (k2 + U[1] * (k3 + U[2] * (int64_t)c))] =
                            prp[c * Lt + m1 + L[0] * (m2 + L[1] * m3)];
                    }
    }
    free(prp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* matvec restatement                                                       */
/* ------------------------------------------------------------------------ */
#define DEF_MATVEC(SUF, CT, RT)                                                              \
    int gor_matvec_##SUF(const int64_t S[3], const int64_t T[3], const int32_t sym[3],        \
                         const CT *fur, const CT *x, CT *y, int64_t nrhs)                     \
    {                                                                                          \
        opr_##SUF op;                                                                          \
        memset(&op, 0, sizeof(op));                                                            \
        if (gor_dims(S, T, sym, op.B, op.U)) return 1;                                         \
        for (int d = 0; d < 3; ++d) { op.S[d] = S[d]; op.T[d] = T[d]; }                        \
        const int64_t Ut = op.U[0] * op.U[1] * op.U[2] * 6;                                    \
        for (int b = 0; b < 8; ++b) op.fur[b] = (CT *)(fur + b * Ut);                          \
        for (int d = 0; d < 3; ++d) {                                                          \
            op.phz[d] = (CT *)malloc(sizeof(CT) * (size_t)op.B[d]);                            \
            for (int64_t k = 0; k < op.B[d]; ++k) {                                            \
                double a = -M_PI * (double)k / (double)op.B[d];                                \
                op.phz[d][k] = (CT)(cos(a) + I * sin(a));                                      \
            }                                                                                  \
        }                                                                                      \
        const int64_t sn = S[0] * S[1] * S[2] * 3, tn = T[0] * T[1] * T[2] * 3;                \
        for (int64_t r = 0; r < nrhs; ++r) {                                                   \
            arr_##SUF a = arr_new_##SUF(S);                                                    \
            memcpy(a.v, x + r * sn, sizeof(CT) * (size_t)sn);                                  \
            arr_##SUF out = brn_##SUF(&op, 0, 0, a);                                           \
            memcpy(y + r * tn, out.v, sizeof(CT) * (size_t)tn);                                \
            free(out.v);                                                                       \
        }                                                                                      \
        for (int d = 0; d < 3; ++d) free(op.phz[d]);                                           \
        return 0;                                                                              \
    }
DEF_MATVEC(c128, double _Complex, double)
DEF_MATVEC(c64, float _Complex, float)

/* plain batched 1-D FFT along one axis, exported so the tests can pin the
 * FFT itself against numpy and the DFT definition */
int gor_fft_axis_c128(zc *v, const int64_t d[3], int64_t ncomp, int32_t dir, int32_t sign)
{
    fft_axis_c128(v, d, ncomp, dir, sign, 1.0);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* brute force: direct circular sum over the real-space circulant           */
/* ------------------------------------------------------------------------ */
/* y_i(t) = sum_s sum_j c_ij[(t - s) mod L] x_j(s), t < T, s < S, with the
 * lower triangle of c taken from the upper one (the matvec only stores the
 * upper triangle).  Independent of any FFT or branch logic. */
int gor_brute_crc(const int64_t S[3], const int64_t T[3], const zc *crc, const zc *x, zc *y)
{
    int64_t L[3];
    for (int d = 0; d < 3; ++d) { if ((S[d] + T[d]) & 1) return 1; L[d] = S[d] + T[d]; }
    const int64_t Sn = S[0] * S[1] * S[2], Tn = T[0] * T[1] * T[2];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < Tn; ++t) {
        int64_t t1 = t % T[0], t2 = (t / T[0]) % T[1], t3 = t / (T[0] * T[1]);
        zc acc[3] = {0, 0, 0};
        for (int64_t s = 0; s < Sn; ++s) {
            int64_t s1 = s % S[0], s2 = (s / S[0]) % S[1], s3 = s / (S[0] * S[1]);
            int64_t n1 = ((t1 - s1) % L[0] + L[0]) % L[0];
            int64_t n2 = ((t2 - s2) % L[1] + L[1]) % L[1];
            int64_t n3 = ((t3 - s3) % L[2] + L[2]) % L[2];
            const zc *c = crc + 9 * (n1 + L[0] * (n2 + L[1] * n3));
            zc xv[3] = {x[s], x[s + Sn], x[s + 2 * Sn]};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    int r = i < j ? i : j, q = i < j ? j : i;
                    acc[i] += c[r + 3 * q] * xv[j];
                }
        }
        y[t] = acc[0]; y[t + Tn] = acc[1]; y[t + 2 * Tn] = acc[2];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* circulant embeddings                                                     */
/* ------------------------------------------------------------------------ */
/* self: c[n] = toe[n] for n < N, 0 at n = N, and toe[2N - n] with the
 * component flips f_i f_j for n > N (egoToeCrc!). toe is (3,3,N1,N2,N3). */
int gor_crc_self(const int64_t N[3], const zc *toe, zc *crc)
{
    const int64_t L[3] = {2 * N[0], 2 * N[1], 2 * N[2]};
    const int64_t Lt = L[0] * L[1] * L[2];
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < Lt; ++p) {
        int64_t n[3] = {p % L[0], (p / L[0]) % L[1], p / (L[0] * L[1])};
        zc *dst = crc + 9 * p;
        if (n[0] == N[0] || n[1] == N[1] || n[2] == N[2]) {
            for (int e = 0; e < 9; ++e) dst[e] = 0;
            continue;
        }
        int64_t q[3];
        double f[3];
        for (int d = 0; d < 3; ++d) {
            if (n[d] < N[d]) { q[d] = n[d]; f[d] = 1.0; }
            else { q[d] = 2 * N[d] - n[d]; f[d] = -1.0; }
        }
        const zc *src = toe + 9 * (q[0] + N[0] * (q[1] + N[1] * q[2]));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                dst[i + 3 * j] = src[i + 3 * j] * (i == j ? 1.0 : f[i] * f[j]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* test Green tensors                                                       */
/* ------------------------------------------------------------------------ */
/* Model 0, point collocation (synthetic, cheap; any size):
 *   G_ij(r) = V (k^2 delta_ij + d_i d_j) [e^{ikr} / (4 pi r zeta^2)],
 *   k = 2 pi zeta, r in wavelengths, V the cell volume, and
 *   G_ij(0) = -delta_ij / (3 zeta^2).
 * Model 1, face-pair surface integrals (the reference's formulation,
 * egoFunInn! / egoSrfAdp! / srfKer / srfSum!, src/glaGen.jl:322-338,
 * 531-598, with cubFac / facPar / srfScl, src/glaInt.jl:271-328, and the
 * scalar kernel sclEgo, src/glaInt.jl:41-44): 36 face-pair integrals of
 * g(|r_t - r_s|) over [0,1]^4, scaled by srfScl and combined with the signs of
 * srfSum!; identity -1/zeta^2 on the diagonal at zero separation
 * (src/glaGen.jl:128-130).  The reference integrates with adaptive hcubature
 * (reltol 1e-6) and replaces the touching-face pairs with DIRECTFN singular
 * rules; here every pair uses a fixed tensor Gauss-Legendre rule (order qf
 * for separated cells, qn for cells that touch), so values near contact agree
 * with the reference only to the quadrature error of that rule.
 * Model 2 is model 1 with the separated-cell order chosen by distance in
 * cells (6 below 3, 4 below 8, 3 below 24, 2 beyond), the rule the library's
 * GPU setup uses. */
static void gl_rule(int n, double *x, double *w)
{
    /* Newton iteration on P_n (textbook Golub-free construction) on [0,1] */
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) { double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * pp * pp);
    }
}

int gor_gl_rule(int32_t n, double *x, double *w) { gl_rule(n, x, w); return 0; }

/* cube faces, (3 coords, 4 corners, 6 faces) like cubFac */
static void cub_fac(const double s[3], double F[6][4][3])
{
    const double a = s[0] / 2, b = s[1] / 2, c = s[2] / 2;
    const double f[6][4][3] = {
        {{-a, -b, -c}, {-a, b, -c}, {-a, b, c}, {-a, -b, c}},   /* yzL */
        {{a, -b, -c}, {a, -b, c}, {a, b, c}, {a, b, -c}},       /* yzU */
        {{-a, -b, -c}, {-a, -b, c}, {a, -b, c}, {a, -b, -c}},   /* xzL */
        {{-a, b, -c}, {a, b, -c}, {a, b, c}, {-a, b, c}},       /* xzU */
        {{-a, -b, -c}, {a, -b, -c}, {a, b, -c}, {-a, b, -c}},   /* xyL */
        {{-a, -b, c}, {-a, b, c}, {a, b, c}, {a, -b, c}}};      /* xyU */
    memcpy(F, f, sizeof(f));
}

static void tensor_facepair(const double sep[3], const double s[3], zc zeta, int qf, int qn, int adaptive,
                            zc out[9])
{
    double F[6][4][3];
    cub_fac(s, F);
    int touch = 1;
    for (int d = 0; d < 3; ++d) if (fabs(sep[d]) > s[d] + 1e-8) touch = 0;
    if (adaptive && !touch) {
        /* order by separation in cells: 6 below 3, 4 below 8, 3 below 24, else 2 */
        double hm = fmax(s[0], fmax(s[1], s[2]));
        double dist = sqrt(sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2]) / hm;
        qf = dist < 3.0 ? 6 : (dist < 8.0 ? 4 : (dist < 24.0 ? 3 : 2));
    }
    /* touching cells: source and target rules of different order so that no
     * source node coincides with a target node on shared faces */
    const int q = touch ? qn : qf, qt = touch ? qn + 1 : qf;
    double *gx = (double *)malloc(sizeof(double) * q), *gw = (double *)malloc(sizeof(double) * q);
    double *hx = (double *)malloc(sizeof(double) * qt), *hw = (double *)malloc(sizeof(double) * qt);
    gl_rule(q, gx, gw);
    gl_rule(qt, hx, hw);
    zc srf[36];
    for (int fp = 0; fp < 36; ++fp) {
        const int ft = fp / 6, fs = fp % 6;    /* facPar: (target face, source face) */
        /* srfScl indexes (srcFId-1)*6+trgFId with srcFId taken from the
         * first facPar column; all cells share one scale here, so the
         * area / length ratio is the same either way round */
        double scale = ((ft < 2) ? s[1] * s[2] : (ft < 4 ? s[0] * s[2] : s[0] * s[1])) /
                       ((fs < 2) ? s[0] : (fs < 4 ? s[1] : s[2]));
        double et1[3], et2[3], es1[3], es2[3];
        for (int d = 0; d < 3; ++d) {
            et1[d] = F[ft][1][d] - F[ft][0][d]; et2[d] = F[ft][3][d] - F[ft][0][d];
            es1[d] = F[fs][1][d] - F[fs][0][d]; es2[d] = F[fs][3][d] - F[fs][0][d];
        }
        zc acc = 0;
        for (int a = 0; a < q; ++a)
            for (int b = 0; b < q; ++b)
                for (int c = 0; c < qt; ++c)
                    for (int e = 0; e < qt; ++e) {
                        double r2 = 0;
                        for (int d = 0; d < 3; ++d) {
                            double dv = (F[ft][0][d] + hx[c] * et1[d] + hx[e] * et2[d]) -
                                        (F[fs][0][d] + gx[a] * es1[d] + gx[b] * es2[d]) + sep[d];
                            r2 += dv * dv;
                        }
                        double r = sqrt(r2);
                        zc g = cexp(2.0 * M_PI * I * r * zeta) / (4.0 * M_PI * r * zeta * zeta);
                        acc += gw[a] * gw[b] * hw[c] * hw[e] * g;
                    }
        srf[fp] = acc * scale;
    }
    free(gx); free(gw); free(hx); free(hw);
    /* srfSum! (1-based srfMat indices shifted by one) */
    zc *m = srf;
    out[0] = m[14] - m[15] - m[20] + m[21] + m[28] - m[29] - m[34] + m[35];
    out[1] = -m[12] + m[13] + m[18] - m[19];
    out[2] = -m[24] + m[25] + m[30] - m[31];
    out[3] = -m[2] + m[3] + m[8] - m[9];
    out[4] = m[0] - m[1] - m[6] + m[7] + m[28] - m[29] - m[34] + m[35];
    out[5] = -m[26] + m[27] + m[32] - m[33];
    out[6] = -m[4] + m[5] + m[10] - m[11];
    out[7] = -m[16] + m[17] + m[22] - m[23];
    out[8] = m[0] - m[1] - m[6] + m[7] + m[14] - m[15] - m[20] + m[21];
    if (sep[0] == 0 && sep[1] == 0 && sep[2] == 0) {
        zc id = 1.0 / (zeta * zeta);
        out[0] -= id; out[4] -= id; out[8] -= id;
    }
}

static void tensor_colloc(const double sep[3], const double s[3], zc zeta, zc out[9])
{
    const double r = sqrt(sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2]);
    for (int e = 0; e < 9; ++e) out[e] = 0;
    if (r < 1e-14) {
        zc v = -1.0 / (3.0 * zeta * zeta);
        out[0] = v; out[4] = v; out[8] = v;
        return;
    }
    const zc k = 2.0 * M_PI * zeta;
    const zc g = cexp(I * k * r) / (4.0 * M_PI * r * zeta * zeta);
    const zc a = k * k + I * k / r - 1.0 / (r * r);
    const zc b = 3.0 / (r * r) - 3.0 * I * k / r - k * k;
    const double V = s[0] * s[1] * s[2];
    const zc pre = V * g;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[i + 3 * j] = pre * ((i == j ? a : 0) + b * sep[i] * sep[j] / (r * r));
}

/* srfSum! (src/glaGen.jl:531-555) with 0-based srfMat indices; out is the
 * column-major 3x3 tensor */
static void srf_sum(const zc *m, zc out[9])
{
    out[0] = m[14] - m[15] - m[20] + m[21] + m[28] - m[29] - m[34] + m[35];
    out[1] = -m[12] + m[13] + m[18] - m[19];
    out[2] = -m[24] + m[25] + m[30] - m[31];
    out[3] = -m[2] + m[3] + m[8] - m[9];
    out[4] = m[0] - m[1] - m[6] + m[7] + m[28] - m[29] - m[34] + m[35];
    out[5] = -m[26] + m[27] + m[32] - m[33];
    out[6] = -m[4] + m[5] + m[10] - m[11];
    out[7] = -m[16] + m[17] + m[22] - m[23];
    out[8] = m[0] - m[1] - m[6] + m[7] + m[14] - m[15] - m[20] + m[21];
}

/* one face pair (target face ft, source face fs, facPar order) over [0,1]^4
 * with an order-q tensor Gauss-Legendre rule, times srfScl: the egoSrfAdp!
 * value (src/glaGen.jl:559-598) with a fixed rule instead of hcubature */
/* target faces FT of a cell of sizes sT, source faces FS of sizes sS; the
 * srfScl factor (src/glaInt.jl:284-305) is indexed (srcFId - 1) * 6 + trgFId
 * while facPar puts the target face first (src/glaInt.jl:271-280), so the
 * factor of pair (ft, fs) is the sS-area of face ft over sT along the normal
 * of fs */
static zc pair_gl2(const double FT[6][4][3], const double FS[6][4][3], int ft, int fs, const double sep[3],
                   const double sT[3], const double sS[3], zc zeta, int q, const double *gx, const double *gw)
{
    const double scale = ((ft < 2) ? sS[1] * sS[2] : (ft < 4 ? sS[0] * sS[2] : sS[0] * sS[1])) /
                         ((fs < 2) ? sT[0] : (fs < 4 ? sT[1] : sT[2]));
    double et1[3], et2[3], es1[3], es2[3];
    for (int d = 0; d < 3; ++d) {
        et1[d] = FT[ft][1][d] - FT[ft][0][d]; et2[d] = FT[ft][3][d] - FT[ft][0][d];
        es1[d] = FS[fs][1][d] - FS[fs][0][d]; es2[d] = FS[fs][3][d] - FS[fs][0][d];
    }
    zc acc = 0;
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b)
            for (int c = 0; c < q; ++c)
                for (int e = 0; e < q; ++e) {
                    double r2 = 0;
                    for (int d = 0; d < 3; ++d) {
                        double dv = (FT[ft][0][d] + gx[c] * et1[d] + gx[e] * et2[d]) -
                                    (FS[fs][0][d] + gx[a] * es1[d] + gx[b] * es2[d]) + sep[d];
                        r2 += dv * dv;
                    }
                    double r = sqrt(r2);
                    zc g = cexp(2.0 * M_PI * I * r * zeta) / (4.0 * M_PI * r * zeta * zeta);
                    acc += gw[a] * gw[b] * gw[c] * gw[e] * g;
                }
    return acc * scale;
}

static zc pair_gl(const double F[6][4][3], int ft, int fs, const double sep[3], const double s[3], zc zeta,
                  int q, const double *gx, const double *gw)
{
    return pair_gl2(F, F, ft, fs, sep, s, s, zeta, q, gx, gw);
}

/* Model 3, the reference's construction: egoFunInn! (all 36 face pairs,
 * src/glaGen.jl:322-338) for cells more than one cell apart on some axis, and
 * egoFunSng! (src/glaGen.jl:350-525) for offsets in {0,1}^3: the face pairs
 * its mask marks take the singular panel values wS / wE / wV divided by the
 * cell volume (src/glaGen.jl:117-121), the others the regular integral.
 * Regular pairs use an order-q tensor Gauss-Legendre rule, q from the
 * separation in cells and the cell size in wavelengths (converged to about
 * 1e-14 relative for every pair; the reference's hcubature stops at 1e-6).
 * A separation with a negative component in {-1, 0, 1} cells takes the value
 * at its absolute offset with the component flips f_i f_j, as the contact
 * path of the reference does through the contact volume's circulant
 * (egoCntOut!, src/glaGen.jl:250-304). */
enum { S1_ = 1, S2_, S3_, E1_, E2_, E3_, E4_, E5_, E6_, E7_, E8_, E9_, V1_, V2_, V3_, V4_, V5_, V6_ };
#define __ 0
/* [offset o1 + 2 o2 + 4 o3][target face][source face], faces yzL yzU xzL
 * xzU xyL xyU: the corVal tables of egoFunSng! (mask = nonzero entry) */
static const unsigned char COR[8][6][6] = {
    /* (0,0,0) */
    {{S1_, __, E7_, E7_, E8_, E8_}, {__, S1_, E7_, E7_, E8_, E8_}, {E7_, E7_, S2_, __, E9_, E9_},
     {E7_, E7_, __, S2_, E9_, E9_}, {E8_, E8_, E9_, E9_, S3_, __}, {E8_, E8_, E9_, E9_, __, S3_}},
    /* (1,0,0) */
    {{__, S1_, E7_, E7_, E8_, E8_}, {__, __, __, __, __, __}, {__, E7_, E3_, __, V6_, V6_},
     {__, E7_, __, E3_, V6_, V6_}, {__, E8_, V6_, V6_, E5_, __}, {__, E8_, V6_, V6_, __, E5_}},
    /* (0,1,0) */
    {{E1_, __, __, E7_, V5_, V5_}, {__, E1_, __, E7_, V5_, V5_}, {E7_, E7_, __, S2_, E9_, E9_},
     {__, __, __, __, __, __}, {V5_, V5_, __, E9_, E6_, __}, {V5_, V5_, __, E9_, __, E6_}},
    /* (1,1,0) */
    {{__, E1_, __, E7_, V5_, V5_}, {__, __, __, __, __, __}, {__, E7_, __, E3_, V6_, V6_},
     {__, __, __, __, __, __}, {__, V5_, __, V6_, V3_, __}, {__, V5_, __, V6_, __, V3_}},
    /* (0,0,1) */
    {{E2_, __, V4_, V4_, __, E8_}, {__, E2_, V4_, V4_, __, E8_}, {V4_, V4_, E4_, __, __, E9_},
     {V4_, V4_, __, E4_, __, E9_}, {E8_, E8_, E9_, E9_, __, S3_}, {__, __, __, __, __, __}},
    /* (1,0,1) */
    {{__, E2_, V4_, V4_, __, E8_}, {__, __, __, __, __, __}, {__, V4_, V2_, __, __, V6_},
     {__, V4_, __, V2_, __, V6_}, {__, E8_, V6_, V6_, __, E5_}, {__, __, __, __, __, __}},
    /* (0,1,1) */
    {{V1_, __, __, V4_, __, V5_}, {__, V1_, __, V4_, __, V5_}, {V4_, V4_, __, E4_, __, E9_},
     {__, __, __, __, __, __}, {V5_, V5_, __, E9_, __, E6_}, {__, __, __, __, __, __}},
    /* (1,1,1) */
    {{__, V1_, __, V4_, __, V5_}, {__, __, __, __, __, __}, {__, V4_, __, V2_, __, V6_},
     {__, __, __, __, __, __}, {__, V5_, __, V6_, __, V3_}, {__, __, __, __, __, __}}};
#undef __

static zc g_w18[18];   /* [wS(3) wE(9) wV(6)] / cell volume, set by gor_set_singular */

void gor_set_singular(const zc *w18) { memcpy(g_w18, w18, sizeof(g_w18)); }

static int ref_order(const double sep[3], const double s[3], zc zeta)
{
    double D = 0, hm = 0;
    for (int d = 0; d < 3; ++d) {
        D = fmax(D, fabs(sep[d]) / s[d]);
        hm = fmax(hm, s[d]);
    }
    int q = D <= 2.5 ? 12 : (D <= 4.5 ? 10 : (D <= 8.5 ? 8 : 6));
    const int qp = 6 + (int)ceil(6.0 * 2.0 * M_PI * cabs(zeta) * hm);
    return q > qp ? q : (qp > 20 ? 20 : qp);
}

static void tensor_ref(const double sep[3], const double s[3], zc zeta, zc out[9])
{
    double F[6][4][3];
    cub_fac(s, F);
    int near = 1, o[3];
    double f[3], asep[3];
    for (int d = 0; d < 3; ++d) {
        const double t = sep[d] / s[d];
        const long r = lround(t);
        if (fabs(t - (double)r) > 1e-9 || labs(r) > 1) near = 0;
        o[d] = (int)labs(r);
        f[d] = r < 0 ? -1.0 : 1.0;
        asep[d] = (double)o[d] * s[d];
    }
    const int q = ref_order(sep, s, zeta);
    double gx[24], gw[24];
    gl_rule(q, gx, gw);
    zc srf[36];
    for (int fp = 0; fp < 36; ++fp) {
        const int ft = fp / 6, fs = fp % 6;
        const int code = near ? COR[o[0] + 2 * o[1] + 4 * o[2]][ft][fs] : 0;
        srf[fp] = code ? g_w18[code - 1] : pair_gl(F, ft, fs, near ? asep : sep, s, zeta, q, gx, gw);
    }
    srf_sum(srf, out);
    if (near)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (i != j) out[i + 3 * j] *= f[i] * f[j];
    if (sep[0] == 0 && sep[1] == 0 && sep[2] == 0) {
        zc id = 1.0 / (zeta * zeta);
        out[0] -= id; out[4] -= id; out[8] -= id;
    }
}

static void tensor_at(int model, const double sep[3], const double s[3], zc zeta, int qf, int qn, zc out[9])
{
    if (model == 3) tensor_ref(sep, s, zeta, out);
    else if (model == 1 || model == 2) tensor_facepair(sep, s, zeta, qf, qn, model == 2, out);
    else tensor_colloc(sep, s, zeta, out);
}

/* Toeplitz tensor over non-negative offsets (3,3,N1,N2,N3) */
int gor_toe_model(int32_t model, const int64_t N[3], const double scl[3], double zr, double zi,
                  int32_t qf, int32_t qn, zc *toe)
{
    const zc zeta = zr + I * zi;
    const int64_t Nt = N[0] * N[1] * N[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < Nt; ++p) {
        double sep[3] = {(double)(p % N[0]) * scl[0], (double)((p / N[0]) % N[1]) * scl[1],
                         (double)(p / (N[0] * N[1])) * scl[2]};
        tensor_at(model, sep, scl, zeta, qf, qn, toe + 9 * p);
    }
    return 0;
}

/* External circulant over separations: along axis d,
 *   sep_d(n) = delta_d + n h_d          for n < T_d,
 *   0-plane                             at n = T_d,
 *   sep_d(n) = delta_d + (n - L_d) h_d  for n > T_d,
 * delta = (first target centre) - (first source centre), one common cell
 * size h (grdSel with sepGrd, src/glaGen.jl:603-643). */
int gor_crc_ext(int32_t model, const int64_t S[3], const int64_t T[3], const double scl[3],
                const double delta[3], double zr, double zi, int32_t qf, int32_t qn, zc *crc)
{
    const zc zeta = zr + I * zi;
    int64_t L[3];
    for (int d = 0; d < 3; ++d) { if ((S[d] + T[d]) & 1) return 1; L[d] = S[d] + T[d]; }
    const int64_t Lt = L[0] * L[1] * L[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < Lt; ++p) {
        int64_t n[3] = {p % L[0], (p / L[0]) % L[1], p / (L[0] * L[1])};
        zc *dst = crc + 9 * p;
        if (n[0] == T[0] || n[1] == T[1] || n[2] == T[2]) {
            for (int e = 0; e < 9; ++e) dst[e] = 0;
            continue;
        }
        double sep[3];
        for (int d = 0; d < 3; ++d)
            sep[d] = delta[d] + (double)(n[d] < T[d] ? n[d] : n[d] - L[d]) * scl[d];
        tensor_at(model, sep, scl, zeta, qf, qn, dst);
    }
    return 0;
}

/* Separated cells of different sizes (a partitioned operator's pair,
 * genEgoExt! with partition volumes, src/glaMemSup.jl:64-89): all 36 face
 * pairs of egoFunOut! (src/glaGen.jl:233-246) between the faces of a target
 * cell of sizes sT and a source cell of sizes sS, srfScl(sT, sS) scaling,
 * srfSum!.  The order follows the separation in mean cell sizes. */
static void tensor_mixed(const double sep[3], const double sT[3], const double sS[3], zc zeta, zc out[9])
{
    double FT[6][4][3], FS[6][4][3], mid[3], hm = 0;
    cub_fac(sT, FT);
    cub_fac(sS, FS);
    for (int d = 0; d < 3; ++d) {
        mid[d] = 0.5 * (sT[d] + sS[d]);
        hm = fmax(hm, fmax(sT[d], sS[d]));
    }
    const int q = ref_order(sep, mid, zeta);
    const int qp = 6 + (int)ceil(6.0 * 2.0 * M_PI * cabs(zeta) * hm);
    const int qq = q > qp ? q : (qp > 20 ? 20 : qp);
    double gx[24], gw[24];
    gl_rule(qq, gx, gw);
    zc srf[36];
    for (int fp = 0; fp < 36; ++fp)
        srf[fp] = pair_gl2(FT, FS, fp / 6, fp % 6, sep, sT, sS, zeta, qq, gx, gw);
    srf_sum(srf, out);
}

/* Circulant of one (source, target) partition pair: as gor_crc_ext with
 * target cells sT, source cells sS and the separation lattice step M (the
 * coarser cell size per axis, the grid step of both partition volumes,
 * sepGrd, src/glaGen.jl:603-629). */
int gor_crc_prt(const int64_t S[3], const int64_t T[3], const double sT[3], const double sS[3], const double M[3],
                const double delta[3], double zr, double zi, zc *crc)
{
    const zc zeta = zr + I * zi;
    int64_t L[3];
    for (int d = 0; d < 3; ++d) { if ((S[d] + T[d]) & 1) return 1; L[d] = S[d] + T[d]; }
    const int64_t Lt = L[0] * L[1] * L[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < Lt; ++p) {
        int64_t n[3] = {p % L[0], (p / L[0]) % L[1], p / (L[0] * L[1])};
        zc *dst = crc + 9 * p;
        if (n[0] == T[0] || n[1] == T[1] || n[2] == T[2]) {
            for (int e = 0; e < 9; ++e) dst[e] = 0;
            continue;
        }
        double sep[3];
        for (int d = 0; d < 3; ++d)
            sep[d] = delta[d] + (double)(n[d] < T[d] ? n[d] : n[d] - L[d]) * M[d];
        tensor_mixed(sep, sT, sS, zeta, dst);
    }
    return 0;
}

/* Model 3 at one separation (3x3, column-major) */
int gor_tensor_ref(const double sep[3], const double s[3], double zr, double zi, zc *out)
{
    tensor_ref(sep, s, zr + I * zi, out);
    return 0;
}

/* The tensor between one target and one source cell (3x3, column-major) */
int gor_tensor_mixed(const double sep[3], const double sT[3], const double sS[3], double zr, double zi, zc *out)
{
    tensor_mixed(sep, sT, sS, zr + I * zi, out);
    return 0;
}

int gor_num_threads(void) { return omp_get_max_threads(); }
void gor_set_threads(int32_t n) { if (n > 0) omp_set_num_threads(n); }
