/*
 * TEST INFRASTRUCTURE ONLY -- part of the CPU oracle, never part of the product.
 *
 * Precision-templated half of the oracle: a plain CPU FFT and a restatement of
 * the reference's host matvec (egoBrnHst! and its split / merge / leaf
 * multiply helpers).  Included twice by gla_oracle.c, once with R = double
 * (the parity oracle) and once with R = float (the ComplexF32 CPU baseline).
 *
 * Macros expected from the includer:
 *   R      real type (double / float)
 *   C      complex type (double _Complex / float _Complex)
 *   SFX(x) name suffixer (x##_c128 / x##_c64)
 *
 * Array convention throughout: Julia column-major, index
 *   (i1, i2, i3, c) -> i1 + n1*(i2 + n2*(i3 + n3*c)),
 * i.e. axis 1 fastest, then axis 2, axis 3, field component (1..3).
 */

/* ------------------------------------------------------------------------ */
/* FFT: iterative mixed-radix Stockham, unnormalised, sign -1 forward.      */
/* Stage t with radix r and running span ns reads x[j + q*n/r], multiplies  */
/* by w_{ns*r}^{q*(j mod ns)}, applies DFT_r, and writes                    */
/* y[(j/ns)*ns*r + (j mod ns) + q*ns].                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int nst;
    int rad[48];
    int ns[48];
    C *tw[48];      /* tw[t][q*ns + k] = exp(-2 pi i q k / (ns r)) */
    C *root[48];    /* generic radices: exp(-2 pi i q / r), q < r     */
} SFX(fplan);

static void SFX(fplan_free)(SFX(fplan) *p)
{
    for (int t = 0; t < p->nst; ++t) { free(p->tw[t]); free(p->root[t]); }
    free(p);
}

static SFX(fplan) *SFX(fplan_make)(int n)
{
    SFX(fplan) *p = (SFX(fplan) *)calloc(1, sizeof(SFX(fplan)));
    p->n = n;
    int m = n;
    /* radix-4 first, then 2, 3, 5, then any remaining prime */
    while (m > 1) {
        int r;
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
        else if (m % 3 == 0) r = 3;
        else if (m % 5 == 0) r = 5;
        else { r = m; for (int f = 7; f * f <= m; f += 2) if (m % f == 0) { r = f; break; } }
        p->rad[p->nst++] = r;
        m /= r;
    }
    int ns = 1;
    for (int t = 0; t < p->nst; ++t) {
        int r = p->rad[t];
        p->ns[t] = ns;
        p->tw[t] = (C *)malloc(sizeof(C) * (size_t)r * ns);
        for (int q = 0; q < r; ++q)
            for (int k = 0; k < ns; ++k) {
                double a = -2.0 * M_PI * (double)q * (double)k / ((double)ns * r);
                p->tw[t][q * ns + k] = (C)(cos(a) + I * sin(a));
            }
        p->root[t] = (C *)malloc(sizeof(C) * r);
        for (int q = 0; q < r; ++q) {
            double a = -2.0 * M_PI * (double)q / (double)r;
            p->root[t][q] = (C)(cos(a) + I * sin(a));
        }
        ns *= r;
    }
    return p;
}

/* in-place on x (length n), work w (length n); sign -1 forward, +1 inverse */
static void SFX(fft_exec)(const SFX(fplan) *p, C *x, C *w, int sign)
{
    const int n = p->n;
    C *src = x, *dst = w;
    C a[64];
    C *ag = NULL;
    for (int t = 0; t < p->nst; ++t) {
        const int r = p->rad[t], ns = p->ns[t], m = n / r;
        const C *tw = p->tw[t];
        C *buf = a;
        if (r > 64) { if (!ag) ag = (C *)malloc(sizeof(C) * (size_t)n); buf = ag; }
        for (int j = 0; j < m; ++j) {
            const int k = j % ns;
            for (int q = 0; q < r; ++q) {
                C v = src[j + q * m];
                C tq = tw[q * ns + k];
                if (sign > 0) tq = conj(tq);
                buf[q] = v * tq;
            }
            const int base = (j / ns) * ns * r + k;
            if (r == 2) {
                dst[base] = buf[0] + buf[1];
                dst[base + ns] = buf[0] - buf[1];
            } else if (r == 4) {
                C s02 = buf[0] + buf[2], d02 = buf[0] - buf[2];
                C s13 = buf[1] + buf[3], d13 = buf[1] - buf[3];
                C jd13 = (sign < 0) ? (-I * d13) : (I * d13);
                dst[base] = s02 + s13;
                dst[base + ns] = d02 + jd13;
                dst[base + 2 * ns] = s02 - s13;
                dst[base + 3 * ns] = d02 - jd13;
            } else if (r == 3) {
                const R c1 = (R)-0.5, s1 = (R)(sign < 0 ? -0.86602540378443864676 : 0.86602540378443864676);
                C s = buf[1] + buf[2], d = buf[1] - buf[2];
                dst[base] = buf[0] + s;
                C u = buf[0] + c1 * s;
                C v = I * s1 * d;
                dst[base + ns] = u + v;
                dst[base + 2 * ns] = u - v;
            } else {
                const C *rt = p->root[t];
                for (int q = 0; q < r; ++q) {
                    C acc = 0;
                    for (int e = 0; e < r; ++e) {
                        C rv = rt[(int)(((long long)e * q) % r)];
                        if (sign > 0) rv = conj(rv);
                        acc += buf[e] * rv;
                    }
                    dst[base + q * ns] = acc;
                }
            }
        }
        C *tmp = src; src = dst; dst = tmp;
    }
    if (src != x) memcpy(x, src, sizeof(C) * (size_t)n);
    free(ag);
}

/* Batched FFT along axis `dir` (0,1,2) of a (d0,d1,d2,ncomp) array, in place.
 * scale multiplies the result (1 for forward, 1/n for the reference's
 * ScaledPlan inverse). */
static void SFX(fft_axis)(C *v, const int64_t d[3], int64_t ncomp, int dir, int sign, R scale)
{
    const int n = (int)d[dir];
    if (n <= 1) { if (scale != (R)1) for (int64_t i = 0; i < d[0] * d[1] * d[2] * ncomp; ++i) v[i] *= scale; return; }
    SFX(fplan) *p = SFX(fplan_make)(n);
    const int64_t str = dir == 0 ? 1 : (dir == 1 ? d[0] : d[0] * d[1]);
    const int64_t o1 = dir == 0 ? d[1] : d[0];
    const int64_t o2 = dir == 2 ? d[1] : d[2];
    const int64_t s1 = dir == 0 ? d[0] : 1;
    const int64_t s2 = dir == 2 ? d[0] : d[0] * d[1];
    const int64_t npen = o1 * o2 * ncomp;
    const int64_t cstr = d[0] * d[1] * d[2];
#pragma omp parallel
    {
        C *buf = (C *)malloc(sizeof(C) * (size_t)n);
        C *wrk = (C *)malloc(sizeof(C) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t pen = 0; pen < npen; ++pen) {
            int64_t a = pen % o1, b = (pen / o1) % o2, c = pen / (o1 * o2);
            C *base = v + c * cstr + a * s1 + b * s2;
            for (int i = 0; i < n; ++i) buf[i] = base[i * str];
            SFX(fft_exec)(p, buf, wrk, sign);
            for (int i = 0; i < n; ++i) base[i * str] = buf[i] * scale;
        }
        free(buf); free(wrk);
    }
    SFX(fplan_free)(p);
}

/* ------------------------------------------------------------------------ */
/* Operator state of the restatement (GlaOprMem, src/glaMem.jl:93-106).     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t S[3], T[3], B[3], U[3];
    C *fur[8];      /* (U1,U2,U3,6): [xx,yy,zz,xy,xz,yz], src/glaMemSup.jl:138-156 */
    C *phz[3];      /* phz_d[k] = exp(-i pi k / B_d), src/glaMemSup.jl:232-235 */
} SFX(opr);

typedef struct { int64_t n[3]; C *v; } SFX(arr);

static SFX(arr) SFX(arr_new)(const int64_t n[3])
{
    SFX(arr) a;
    a.n[0] = n[0]; a.n[1] = n[1]; a.n[2] = n[2];
    a.v = (C *)calloc((size_t)(n[0] * n[1] * n[2] * 3), sizeof(C));
    return a;
}

#define IDX4(a, i, j, k, c) ((i) + (a).n[0] * ((j) + (a).n[1] * ((k) + (a).n[2] * (int64_t)(c))))

/* same-size split: prg = phz[i_dir] * org  (src/glaActSup.jl:72-82) */
static void SFX(spt_same)(SFX(arr) *prg, const SFX(arr) *org, int dir, const C *phz)
{
    const int64_t *n = org->n;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t c = 0; c < 3; ++c)
        for (int64_t k = 0; k < n[2]; ++k)
            for (int64_t j = 0; j < n[1]; ++j)
                for (int64_t i = 0; i < n[0]; ++i) {
                    int64_t id = dir == 0 ? i : (dir == 1 ? j : k);
                    int64_t o = IDX4(*org, i, j, k, c);
                    prg->v[o] = phz[id] * org->v[o];
                }
}

/* resizing split along dir: source length S -> branch length B
 * (src/glaActSup.jl:163-217): rows n < S-B get the spill-over sums,
 * rows up to min(S,B) a plain copy, the rest zero. */
static void SFX(spt_resize)(SFX(arr) *eve, SFX(arr) *odd, const SFX(arr) *org, int dir,
                            int64_t Bd, const C *phz)
{
    const int64_t Sd = org->n[dir];
    const int64_t ovr = Sd > Bd ? Sd - Bd : 0;       /* ovrRng = 1:(S-B) */
    const int64_t stdEnd = Sd < Bd ? Sd : Bd;         /* stdRng = (ovr+1):min(S,B) */
    const int64_t *n = eve->n;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t c = 0; c < 3; ++c)
        for (int64_t k = 0; k < n[2]; ++k)
            for (int64_t j = 0; j < n[1]; ++j)
                for (int64_t i = 0; i < n[0]; ++i) {
                    int64_t id = dir == 0 ? i : (dir == 1 ? j : k);
                    int64_t o = IDX4(*eve, i, j, k, c);
                    C e, d;
                    if (id < ovr) {
                        int64_t i2 = i + (dir == 0) * Bd, j2 = j + (dir == 1) * Bd, k2 = k + (dir == 2) * Bd;
                        C a = org->v[IDX4(*org, i, j, k, c)], b = org->v[IDX4(*org, i2, j2, k2, c)];
                        e = a + b; d = phz[id] * (a - b);
                    } else if (id < stdEnd) {
                        C a = org->v[IDX4(*org, i, j, k, c)];
                        e = a; d = phz[id] * a;
                    } else {
                        e = 0; d = 0;
                    }
                    eve->v[o] = e; odd->v[o] = d;
                }
}

/* same-size merge: eve = 1/2 (eve + conj(phz) odd)  (src/glaActSup.jl:425-436) */
static void SFX(mrg_same)(SFX(arr) *eve, const SFX(arr) *odd, int dir, const C *phz)
{
    const int64_t *n = eve->n;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t c = 0; c < 3; ++c)
        for (int64_t k = 0; k < n[2]; ++k)
            for (int64_t j = 0; j < n[1]; ++j)
                for (int64_t i = 0; i < n[0]; ++i) {
                    int64_t id = dir == 0 ? i : (dir == 1 ? j : k);
                    int64_t o = IDX4(*eve, i, j, k, c);
                    eve->v[o] = (R)0.5 * (eve->v[o] + conj(phz[id]) * odd->v[o]);
                }
}

/* resizing merge onto the target length (src/glaActSup.jl:504-535):
 * top rows n < min(T,B) get 1/2 (e + conj(phz) o), the rows after them
 * 1/2 (e - conj(phz) o). */
static void SFX(mrg_resize)(SFX(arr) *out, const SFX(arr) *eve, const SFX(arr) *odd, int dir,
                            const C *phz)
{
    const int64_t Td = out->n[dir], Bd = eve->n[dir];
    const int64_t top = Td < Bd ? Td : Bd;
    const int64_t *n = out->n;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t c = 0; c < 3; ++c)
        for (int64_t k = 0; k < n[2]; ++k)
            for (int64_t j = 0; j < n[1]; ++j)
                for (int64_t i = 0; i < n[0]; ++i) {
                    int64_t id = dir == 0 ? i : (dir == 1 ? j : k);
                    int64_t src = id < top ? id : id - top;
                    int64_t i2 = dir == 0 ? src : i, j2 = dir == 1 ? src : j, k2 = dir == 2 ? src : k;
                    int64_t so = IDX4(*eve, i2, j2, k2, c);
                    C po = conj(phz[src]) * odd->v[so];
                    out->v[IDX4(*out, i, j, k, c)] =
                        (R)0.5 * (id < top ? (eve->v[so] + po) : (eve->v[so] - po));
                }
}

/* leaf multiply over the eight forward/reverse octants
 * (mulBrnHst! src/glaActSup.jl:671-766, kernel mulKerHst! :879-899). */
static void SFX(mul_brn)(const SFX(opr) *op, int bId, SFX(arr) *prd, const SFX(arr) *org)
{
    const int64_t *B = op->B, *U = op->U;
    int brnSym[3];
    for (int d = 0; d < 3; ++d) brnSym[d] = (bId >> (2 - d)) & 1;
    const C *mod = op->fur[bId];
    for (int div = 0; div < 8; ++div) {
        int64_t mBeg[3], mLen[3], vBeg[3], vStp[3];
        R sgn[3];
        int empty = 0;
        for (int d = 0; d < 3; ++d) {
            int rev = (div >> d) & 1;
            if (U[d] == B[d]) {
                if (!rev) { mBeg[d] = 0; mLen[d] = B[d]; vBeg[d] = 0; vStp[d] = 1; sgn[d] = 1; }
                else { mLen[d] = 0; mBeg[d] = 0; vBeg[d] = 0; vStp[d] = 1; sgn[d] = 0; }
            } else {
                const int64_t T = op->T[d], S = op->S[d];
                int64_t beg, end;   /* 1-based inclusive, as in the reference */
                if (brnSym[d] == 0) {
                    if (!rev) { beg = 1; end = (T + 1) / 2 + (T % 2 == 0); }
                    else { beg = 2; end = S / 2 + 1 - (T % 2 == 0); }
                } else {
                    beg = 1;
                    end = !rev ? (T + 1) / 2 : S / 2;
                }
                mBeg[d] = beg - 1;
                mLen[d] = end - beg + 1;
                if (mLen[d] < 0) mLen[d] = 0;
                if (!rev) { vBeg[d] = 0; vStp[d] = 1; }
                else { vBeg[d] = B[d] - 1; vStp[d] = -1; }
                sgn[d] = (R)(1 - 2 * rev);
            }
            if (mLen[d] <= 0) empty = 1;
        }
        if (empty) continue;
        const R ds0 = sgn[0] * sgn[1], ds1 = sgn[0] * sgn[2], ds2 = sgn[1] * sgn[2];
        const int64_t Ut = U[0] * U[1] * U[2];
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t c = 0; c < mLen[2]; ++c)
            for (int64_t b = 0; b < mLen[1]; ++b)
                for (int64_t a = 0; a < mLen[0]; ++a) {
                    int64_t mi = (mBeg[0] + a) + U[0] * ((mBeg[1] + b) + U[1] * (mBeg[2] + c));
                    int64_t vi0 = vBeg[0] + vStp[0] * a, vi1 = vBeg[1] + vStp[1] * b, vi2 = vBeg[2] + vStp[2] * c;
                    C gxx = mod[mi], gyy = mod[mi + Ut], gzz = mod[mi + 2 * Ut];
                    C gxy = mod[mi + 3 * Ut], gxz = mod[mi + 4 * Ut], gyz = mod[mi + 5 * Ut];
                    C x0 = org->v[IDX4(*org, vi0, vi1, vi2, 0)];
                    C x1 = org->v[IDX4(*org, vi0, vi1, vi2, 1)];
                    C x2 = org->v[IDX4(*org, vi0, vi1, vi2, 2)];
                    prd->v[IDX4(*prd, vi0, vi1, vi2, 0)] += gxx * x0 + ds0 * gxy * x1 + ds1 * gxz * x2;
                    prd->v[IDX4(*prd, vi0, vi1, vi2, 1)] += ds0 * gxy * x0 + gyy * x1 + ds2 * gyz * x2;
                    prd->v[IDX4(*prd, vi0, vi1, vi2, 2)] += ds1 * gxz * x0 + ds2 * gyz * x1 + gzz * x2;
                }
    }
}

/* recursive branch driver: restatement of egoBrnHst! (src/glaAct.jl:63-183)
 * for one source / target partition.  Consumes `org` (frees it) and returns
 * a freshly allocated result, as the reference does with its input. */
static SFX(arr) SFX(brn)(const SFX(opr) *op, int lvl, int bId, SFX(arr) org)
{
    const int64_t *B = op->B;
    if (lvl > 0) SFX(fft_axis)(org.v, org.n, 3, lvl - 1, -1, (R)1);
    SFX(arr) ret;
    if (lvl < 3) {
        const int d = lvl;   /* split direction lvl+1 (1-based) */
        SFX(arr) eve, odd;
        if (org.n[d] == B[d]) {
            odd = SFX(arr_new)(org.n);
            SFX(spt_same)(&odd, &org, d, op->phz[d]);
            eve = org;
        } else {
            int64_t pn[3] = {org.n[0], org.n[1], org.n[2]};
            pn[d] = B[d];
            eve = SFX(arr_new)(pn);
            odd = SFX(arr_new)(pn);
            SFX(spt_resize)(&eve, &odd, &org, d, B[d], op->phz[d]);
            free(org.v);
        }
        SFX(arr) eveR = SFX(brn)(op, lvl + 1, bId, eve);
        SFX(arr) oddR = SFX(brn)(op, lvl + 1, bId + (1 << (3 - (lvl + 1))), odd);
        if (eveR.n[d] == op->T[d]) {
            SFX(mrg_same)(&eveR, &oddR, d, op->phz[d]);
            free(oddR.v);
            ret = eveR;
        } else {
            int64_t mn[3] = {eveR.n[0], eveR.n[1], eveR.n[2]};
            mn[d] = op->T[d];
            ret = SFX(arr_new)(mn);
            SFX(mrg_resize)(&ret, &eveR, &oddR, d, op->phz[d]);
            free(eveR.v); free(oddR.v);
        }
    } else {
        ret = SFX(arr_new)(B);
        SFX(mul_brn)(op, bId, &ret, &org);
        free(org.v);
    }
    if (lvl > 0) SFX(fft_axis)(ret.v, ret.n, 3, lvl - 1, +1, (R)1 / (R)ret.n[lvl - 1]);
    return ret;
}
