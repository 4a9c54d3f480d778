// Operator construction from geometry on the GPU, in FP64:
//   1. the Green tensor per cell separation by face-pair surface integrals
//      (the reference's formulation: egoFunInn! / egoSrfAdp! / srfKer /
//      srfSum!, src/glaGen.jl:322-338, 531-598; cube faces and scalings of
//      cubFac / srfScl, src/glaInt.jl:284-328; scalar kernel sclEgo,
//      src/glaInt.jl:41-44), with fixed tensor Gauss-Legendre rules whose
//      order falls with the separation instead of adaptive cubature;
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

constexpr int QMAX = 16;   // largest Gauss-Legendre order in the table

struct GreenCfg {
    double scl[3];          // cell size per axis (wavelengths)
    double zr, zi;          // frqPhz
    int qnear;              // order for touching cells (source; target uses qnear + 1)
};

__device__ __forceinline__ void cexp_i(double re, double im, double& o_re, double& o_im)
{
    // exp(i (re + i im)) = exp(-im) (cos re + i sin re)
    double s, c;
    sincos(re, &s, &c);
    const double m = exp(-im);
    o_re = m * c;
    o_im = m * s;
}

// integral over the unit square pair of g(|r_t - r_s + sep|), faces given by
// normal axis (nt, ns) and side (st, ss: 0 lower, 1 upper)
__device__ void face_pair(const GreenCfg& G, const double sep[3], int nt, int st, int ns, int ss, int qs, int qt,
                          const double* __restrict__ gl, double& acc_re, double& acc_im)
{
    // node/weight tables: gl[(q*(q-1)/2 + i)*2 + {0,1}] for order q, node i
    const double* xs = gl + (qs * (qs - 1) / 2) * 2;
    const double* xt = gl + (qt * (qt - 1) / 2) * 2;
    // source face: coordinate ns fixed, the other two spanned
    int sa = (ns + 1) % 3, sb = (ns + 2) % 3;
    int ta = (nt + 1) % 3, tb = (nt + 2) % 3;
    const double h[3] = {G.scl[0], G.scl[1], G.scl[2]};
    const double zr = G.zr, zi = G.zi;
    // 2 pi zeta and 4 pi zeta^2
    const double kr = 2.0 * M_PI * zr, ki = 2.0 * M_PI * zi;
    const double z2r = zr * zr - zi * zi, z2i = 2.0 * zr * zi;
    double ar = 0, ai = 0;
    for (int i1 = 0; i1 < qs; ++i1)
        for (int i2 = 0; i2 < qs; ++i2) {
            double ps[3];
            ps[ns] = (ss ? 0.5 : -0.5) * h[ns];
            ps[sa] = (xs[2 * i1] - 0.5) * h[sa];
            ps[sb] = (xs[2 * i2] - 0.5) * h[sb];
            const double ws = xs[2 * i1 + 1] * xs[2 * i2 + 1];
            for (int j1 = 0; j1 < qt; ++j1)
                for (int j2 = 0; j2 < qt; ++j2) {
                    double pt[3];
                    pt[nt] = (st ? 0.5 : -0.5) * h[nt];
                    pt[ta] = (xt[2 * j1] - 0.5) * h[ta];
                    pt[tb] = (xt[2 * j2] - 0.5) * h[tb];
                    const double dx = pt[0] - ps[0] + sep[0], dy = pt[1] - ps[1] + sep[1], dz = pt[2] - ps[2] + sep[2];
                    const double d = sqrt(dx * dx + dy * dy + dz * dz);
                    // g = exp(i 2 pi d zeta) / (4 pi d zeta^2)
                    double er, ei;
                    cexp_i(kr * d, ki * d, er, ei);
                    const double dr = 4.0 * M_PI * d * z2r, di = 4.0 * M_PI * d * z2i;
                    const double inv = 1.0 / (dr * dr + di * di);
                    const double gr = (er * dr + ei * di) * inv, gi = (ei * dr - er * di) * inv;
                    const double w = ws * xt[2 * j1 + 1] * xt[2 * j2 + 1];
                    ar += w * gr;
                    ai += w * gi;
                }
        }
    acc_re = ar;
    acc_im = ai;
}

// six upper-triangle components at separation sep; srfSum! combination
__device__ void green_at(const GreenCfg& G, const double sep[3], const double* __restrict__ gl, double out[12])
{
    const double* h = G.scl;
    bool touch = true;
    for (int d = 0; d < 3; ++d)
        if (fabs(sep[d]) > h[d] + 1e-8) touch = false;
    const double hm = fmax(h[0], fmax(h[1], h[2]));
    const double dist = sqrt(sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2]) / hm;
    int qs, qt;
    if (touch) { qs = G.qnear; qt = G.qnear + 1; }
    else if (dist < 3.0) { qs = qt = 6; }
    else if (dist < 8.0) { qs = qt = 4; }
    else if (dist < 24.0) { qs = qt = 3; }
    else { qs = qt = 2; }
    // A[i][j]: sum over sides of sign * scale * integral, target normal i,
    // source normal j; sign s_t s_s (lower +1, upper -1), negated off-diagonal
    double A[3][3][2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j][0] = A[i][j][1] = 0;
    const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int p = 0; p < 6; ++p) {
        const int nt = pairs[p][0], ns = pairs[p][1];
        // srfScl: area of the target-normal face over the cell length along
        // the source normal
        const double area = h[(nt + 1) % 3] * h[(nt + 2) % 3];
        const double scale = area / h[ns];
        for (int st = 0; st < 2; ++st)
            for (int ss = 0; ss < 2; ++ss) {
                double r, im;
                face_pair(G, sep, nt, st, ns, ss, qs, qt, gl, r, im);
                const double sg = ((st == ss) ? 1.0 : -1.0) * (nt == ns ? 1.0 : -1.0) * scale;
                A[nt][ns][0] += sg * r;
                A[nt][ns][1] += sg * im;
            }
    }
    // xx = A_yy + A_zz, yy = A_xx + A_zz, zz = A_xx + A_yy; xy, xz, yz
    out[0] = A[1][1][0] + A[2][2][0];  out[1] = A[1][1][1] + A[2][2][1];
    out[2] = A[0][0][0] + A[2][2][0];  out[3] = A[0][0][1] + A[2][2][1];
    out[4] = A[0][0][0] + A[1][1][0];  out[5] = A[0][0][1] + A[1][1][1];
    out[6] = A[0][1][0];  out[7] = A[0][1][1];
    out[8] = A[0][2][0];  out[9] = A[0][2][1];
    out[10] = A[1][2][0]; out[11] = A[1][2][1];
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
                           const double* __restrict__ gl)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)N1 * N2 * N3;
    if (p >= n) return;
    const int a = (int)(p % N1), b = (int)((p / N1) % N2), c = (int)(p / ((int64_t)N1 * N2));
    const double sep[3] = {a * G.scl[0], b * G.scl[1], c * G.scl[2]};
    double o[12];
    green_at(G, sep, gl, o);
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

// external circulant over target-minus-source separations
// (zero for T <= n <= L - S on an axis; L = S + T is the reference's grid)
__global__ void k_crc_ext(double2* __restrict__ crc, int T1, int T2, int T3, int S1, int S2, int S3, int L1, int L2,
                          int L3, double d1, double d2, double d3, GreenCfg G, const double* __restrict__ gl)
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
    green_at(G, sep, gl, o);
    for (int k = 0; k < 6; ++k) crc[k * Lt + p] = make_double2(o[2 * k], o[2 * k + 1]);
}

// branch extraction with truncation: fur[((b*6 + c)*U3 + j3)*U2 + j2)*U1 + j1]
// from the split-transformed grid W (L1,L2,L3, comp6), parity-major per axis;
// a branch shorter than its stored extent repeats its last entry
__global__ void k_gather(double2* __restrict__ fur, const double2* __restrict__ W, int B1, int B2, int B3, int U1,
                         int U2, int U3)
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
    fur[p] = W[c * Lt + m1 + L1 * (m2 + L2 * (int64_t)m3)];
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
    for (int a = 0; a < 3; ++a) G.scl[a] = rat(trg.scl_num[a], trg.scl_den[a]);
    G.zr = opts.frq_phz_re;
    G.zi = opts.frq_phz_im;
    G.qnear = std::max(2, std::min(QMAX - 1, opts.int_ord > 0 ? opts.int_ord / 3 : 10));
    if (G.zr == 0 && G.zi == 0) {
        err = "frqPhz must be nonzero";
        return GLA_ERR_ARG;
    }
    if (!d.self_opr) {
        // overlap check of genEgoCrcExt! (src/glaGen.jl:64-79): edges of each
        // volume along every axis
        double sl[3], su[3], tl[3], tu[3];
        for (int a = 0; a < 3; ++a) {
            const double h = G.scl[a];
            const double so = rat(src.org_num[a], src.org_den[a]), to = rat(trg.org_num[a], trg.org_den[a]);
            sl[a] = so - 0.5 * h * src.cel[a];
            su[a] = so + 0.5 * h * src.cel[a];
            tl[a] = to - 0.5 * h * trg.cel[a];
            tu[a] = to + 0.5 * h * trg.cel[a];
        }
        bool overlap = true, aligned = true;
        for (int a = 0; a < 3; ++a) {
            if (!(tl[a] < su[a] - 1e-12 && sl[a] < tu[a] - 1e-12)) overlap = false;
            const double off = (tl[a] - sl[a]) / G.scl[a];
            if (std::fabs(off - std::round(off)) > 1e-9) aligned = false;
        }
        // volumes whose cell grids coincide may overlap: the operator is then
        // the self operator of the covering grid restricted to the two volumes
        if (overlap && !aligned) {
            err = "Source and target volumes are overlapping.";
            return GLA_ERR_OVERLAP;
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
    DevBuf dgl, dcrc, dw1, dw2, dtab, dtoe, dfur, dcnt;
    SETUP_HIP(hipMalloc(&dgl.p, gl.size() * sizeof(double)));
    SETUP_HIP(hipMemcpy(dgl.p, gl.data(), gl.size() * sizeof(double), hipMemcpyHostToDevice));
    SETUP_HIP(hipMalloc(&dcrc.p, (size_t)Lt * 6 * sizeof(double2)));
    const int thr = 128;
    if (d.self_opr) {
        const int64_t Nt = d.trg_cel[0] * d.trg_cel[1] * d.trg_cel[2];
        SETUP_HIP(hipMalloc(&dtoe.p, (size_t)Nt * 6 * sizeof(double2)));
        hipLaunchKernelGGL(k_toeplitz, dim3((unsigned)((Nt + thr - 1) / thr)), dim3(thr), 0, nullptr, (double2*)dtoe.p,
                           (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2], G, (const double*)dgl.p);
        SETUP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_embed_self, dim3((unsigned)((Lt + 255) / 256)), dim3(256), 0, nullptr, (double2*)dcrc.p,
                           (const double2*)dtoe.p, (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2], L[0], L[1],
                           L[2]);
        SETUP_HIP(hipGetLastError());
    } else {
        // separation of the first target cell centre from the first source one
        double del[3];
        for (int a = 0; a < 3; ++a) {
            const double h = G.scl[a];
            const double t0 = rat(trg.org_num[a], trg.org_den[a]) - h * (trg.cel[a] / 2 - (trg.cel[a] % 2 == 0 ? 0.5 : 0));
            const double s0 = rat(src.org_num[a], src.org_den[a]) - h * (src.cel[a] / 2 - (src.cel[a] % 2 == 0 ? 0.5 : 0));
            del[a] = t0 - s0;
        }
        hipLaunchKernelGGL(k_crc_ext, dim3((unsigned)((Lt + thr - 1) / thr)), dim3(thr), 0, nullptr, (double2*)dcrc.p,
                           (int)d.trg_cel[0], (int)d.trg_cel[1], (int)d.trg_cel[2], (int)d.src_cel[0],
                           (int)d.src_cel[1], (int)d.src_cel[2], L[0], L[1], L[2], del[0], del[1], del[2], G,
                           (const double*)dgl.p);
        SETUP_HIP(hipGetLastError());
    }
    // forward transform over the full L grid: S = L on every axis
    Geom g{};
    for (int a = 0; a < 3; ++a) { g.S[a] = L[a]; g.T[a] = L[a]; g.B[a] = B[a]; g.L[a] = L[a]; }
    g.Y1 = L[1];
    g.Z = L[2];
    g.xs = Lt * 3;
    g.ys = Lt * 3;
    g.w1s = Lt * 3;
    g.w2s = Lt * 3;
    std::vector<double2> tab;
    int64_t off[3][2];
    for (int a = 0; a < 3; ++a) {
        off[a][0] = (int64_t)tab.size();
        for (int n = 0; n < B[a]; ++n) tab.push_back(make_double2(std::cos(-M_PI * n / B[a]), std::sin(-M_PI * n / B[a])));
        off[a][1] = (int64_t)tab.size();
        for (int n = 0; n < B[a]; ++n)
            tab.push_back(make_double2(std::cos(-2 * M_PI * n / B[a]), std::sin(-2 * M_PI * n / B[a])));
    }
    SETUP_HIP(hipMalloc(&dtab.p, tab.size() * sizeof(double2)));
    SETUP_HIP(hipMemcpy(dtab.p, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice));
    DevTab tb{};
    for (int a = 0; a < 3; ++a) {
        tb.phz[a] = (const double2*)dtab.p + off[a][0];
        tb.tw[a] = (const double2*)dtab.p + off[a][1];
    }
    SETUP_HIP(hipMalloc(&dw1.p, (size_t)Lt * 6 * sizeof(double2)));
    SETUP_HIP(hipMalloc(&dw2.p, (size_t)Lt * 6 * sizeof(double2)));
    SETUP_HIP(run_fwd3<double>(g, tb, dcrc.p, dw1.p, dw2.p, 2, nullptr));
    const int64_t Ut = d.fur_ext[0] * d.fur_ext[1] * d.fur_ext[2];
    SETUP_HIP(hipMalloc(&dfur.p, (size_t)Ut * 48 * sizeof(double2)));
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((Ut * 48 + 255) / 256)), dim3(256), 0, nullptr, (double2*)dfur.p,
                       (const double2*)dw2.p, B[0], B[1], B[2], (int)d.fur_ext[0], (int)d.fur_ext[1],
                       (int)d.fur_ext[2]);
    SETUP_HIP(hipGetLastError());
    SETUP_HIP(hipMalloc(&dcnt.p, sizeof(int)));
    SETUP_HIP(hipMemset(dcnt.p, 0, sizeof(int)));
    hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)((Lt * 6 + 255) / 256)), dim3(256), 0, nullptr,
                       (const double2*)dcrc.p, Lt * 6, (int*)dcnt.p);
    SETUP_HIP(hipGetLastError());
    int bad = 0;
    SETUP_HIP(hipMemcpy(&bad, dcnt.p, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) {
        err = "Computed circulant contains non-numeric values.";
        return GLA_ERR_NONFINITE;
    }
    fur.resize((size_t)(Ut * 48));
    SETUP_HIP(hipMemcpy(fur.data(), dfur.p, fur.size() * sizeof(double2), hipMemcpyDeviceToHost));
    return GLA_OK;
#undef SETUP_HIP
}

}  // namespace gla
