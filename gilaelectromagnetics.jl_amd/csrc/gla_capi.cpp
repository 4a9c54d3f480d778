// C ABI of the Green-function operator (include/gila_amd.h): operator state,
// Fourier Green data upload, workspaces and the matvec launch sequence.
#include "gila_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gla_krylov.hpp"
#include "gla_launch.hpp"
#include "gla_setup.hpp"

namespace {

thread_local std::string g_err;

gla_status fail(gla_status s, const std::string& m)
{
    g_err = m;
    return s;
}

#define GLA_HIP(call)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(GLA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int d)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DevGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Green data and tables on the device; shared between an operator and its
// adjoint wrapper.
struct DevData {
    int device = 0;
    void* gf = nullptr;
    void* tab = nullptr;
    void* order = nullptr;
    ~DevData()
    {
        if (gf || tab || order) {
            DevGuard g(device);
            if (gf) (void)hipFree(gf);
            if (tab) (void)hipFree(tab);
            if (order) (void)hipFree(order);
        }
    }
};

bool vol_eq(const gla_vol& a, const gla_vol& b)
{
    for (int d = 0; d < 3; ++d) {
        if (a.cel[d] != b.cel[d]) return false;
        if (a.scl_num[d] * b.scl_den[d] != b.scl_num[d] * a.scl_den[d]) return false;
        if (a.org_num[d] * b.org_den[d] != b.org_num[d] * a.org_den[d]) return false;
    }
    return true;
}

// Work planes of the allocation-free matvec (gla_opr_apply_noalloc): one copy
// of this array per device, held from the first gla_init on that device until
// the process ends.  Its host-side name is not a device address: host code
// uses the address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = (size_t)GLA_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on `device`, looked up once per device with
// that device current
gla_status scratch_addr(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(GLA_ERR_NO_DEVICE, "HIP device " + std::to_string(device) + " is beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        DevGuard g(device);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GLA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return GLA_OK;
}

gla_status check_vol(const gla_vol& v, const char* what)
{
    for (int d = 0; d < 3; ++d) {
        if (v.cel[d] < 1) return fail(GLA_ERR_ARG, std::string(what) + ": cell counts must be positive");
        if (v.scl_den[d] == 0 || v.org_den[d] == 0 || v.scl_num[d] <= 0)
            return fail(GLA_ERR_ARG, std::string(what) + ": cell size must be a positive rational");
    }
    return GLA_OK;
}

}  // namespace

struct gla_opr_s {
    gla_opts opts{};
    gla_vol trg{}, src{};
    gla_dims dims{};
    gla::Geom geom{};
    gla::DevTab tab{};
    std::shared_ptr<DevData> dev;
    std::shared_ptr<std::vector<char>> fur_host;   // reference layout, operator dtype
    void* w1 = nullptr;
    void* w2 = nullptr;
    int64_t ws_nrhs = 0;
    int batches = 0;   // recorded batches that use w1 / w2 (they must not move)
    // device copies of the host arrays of gla_opr_apply_host (grown as
    // needed, kept until release), under their own lock
    void* hx = nullptr;
    void* hy = nullptr;
    int64_t hx_n = 0, hy_n = 0;
    std::mutex hmu;
    std::mutex mu;
    ~gla_opr_s()
    {
        if (w1 || w2 || hx || hy) {
            DevGuard g(opts.device);
            if (w1) (void)hipFree(w1);
            if (w2) (void)hipFree(w2);
            if (hx) (void)hipFree(hx);
            if (hy) (void)hipFree(hy);
        }
    }
};

extern "C" {

const char* gla_last_error(void) { return g_err.c_str(); }

gla_status gla_init(int32_t device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(GLA_ERR_NO_DEVICE, "no HIP device is available");
    if (device < 0 || device >= n)
        return fail(GLA_ERR_NO_DEVICE, "HIP device " + std::to_string(device) + " does not exist");
    void* p = nullptr;
    return scratch_addr(device, &p);
}

}  // extern "C"

namespace {

// glaVolEveGen (src/glaMem.jl:152-176): odd axes get one more cell of a
// proportionally smaller size, the extent unchanged
gla_vol even_gen(const gla_vol& v)
{
    gla_vol o = v;
    for (int a = 0; a < 3; ++a) {
        if (v.cel[a] % 2 == 0) continue;
        const int64_t n = v.cel[a] + 1;
        __int128 num = (__int128)v.scl_num[a] * v.cel[a], den = (__int128)v.scl_den[a] * n;
        __int128 x = num < 0 ? -num : num, y = den < 0 ? -den : den;
        while (y) { const __int128 t = x % y; x = y; y = t; }
        if (x > 1) { num /= x; den /= x; }
        o.cel[a] = n;
        o.scl_num[a] = (int64_t)num;
        o.scl_den[a] = (int64_t)den;
    }
    return o;
}

// an external pair whose cells sum to an odd number on some axis is
// regenerated on both sides, as GlaOprMem does (src/glaMemSup.jl:34-40)
void even_pair(const gla_vol* trg, const gla_vol* src, gla_vol& t2, gla_vol& s2)
{
    t2 = *trg;
    s2 = src ? *src : *trg;
    if (!src || vol_eq(*trg, *src)) return;
    bool odd = false;
    for (int a = 0; a < 3; ++a) odd = odd || (trg->cel[a] + src->cel[a]) % 2 != 0;
    if (odd) {
        t2 = even_gen(*trg);
        s2 = even_gen(*src);
    }
}

// the integer q with a == q * b for positive rationals a, b, or 0
int64_t int_ratio(int64_t an, int64_t ad, int64_t bn, int64_t bd)
{
    const __int128 num = (__int128)an * bd, den = (__int128)ad * bn;
    if (den <= 0 || num <= 0 || num % den != 0) return 0;
    const __int128 q = num / den;
    return q > (__int128)(1 << 20) ? 0 : (int64_t)q;
}

// Device branch lengths by layout:
//   DIMS_FAST  operators built from geometry: a B without a fast transform
//              (2^a x {1, 3, 5}) is embedded into the next fast one -- the
//              circulant only needs L >= S + T - 1, and the lengths with a
//              factor beyond 5 (29, 81, 145) transform several times more
//              slowly;
//   DIMS_EXACT Fourier data handed in: B itself when it has a compiled
//              transform, else the next fast length (the data is then
//              re-embedded, gla::refur);
//   DIMS_REF   the reference's own layout of that data: B on every axis.
// trg / src are the volumes after regeneration (even_pair).
enum { DIMS_FAST = 0, DIMS_EXACT = 1, DIMS_REF = 2 };
// layout of operators built from geometry.  DIMS_FAST measured against
// DIMS_EXACT on the 256x34x34-in-256x256x128 pair (B = (256,145,81)):
// ComplexF32 matvec 0.700 ms on (256,160,96) against 0.922 ms on the exact
// lengths, ComplexF64 1.426 against 1.575 ms (the 145-point axis-2 passes
// dominate).
#ifndef GLA_CREATE_DIMS
#define GLA_CREATE_DIMS DIMS_FAST
#endif

gla_status dims_impl(const gla_vol* trg, const gla_vol* src, int mode, gla_dims* out)
{
    if (!trg || !out) return fail(GLA_ERR_ARG, "null argument");
    gla_status s = check_vol(*trg, "target volume");
    if (s) return s;
    const bool self = (src == nullptr) || vol_eq(*trg, *src);
    const gla_vol& sv = self ? *trg : *src;
    if (!self && (s = check_vol(sv, "source volume"))) return s;
    gla_dims d{};
    d.self_opr = self ? 1 : 0;
    // partitions of GlaExtInf (src/glaMem.jl:190-216): per axis the coarser
    // scale is the lattice (lcm) and the finer volume splits into
    // ratio-strided partitions
    for (int a = 0; a < 3; ++a) {
        const int64_t st = int_ratio(trg->scl_num[a], trg->scl_den[a], sv.scl_num[a], sv.scl_den[a]);
        const int64_t ts = int_ratio(sv.scl_num[a], sv.scl_den[a], trg->scl_num[a], trg->scl_den[a]);
        if (st == 0 && ts == 0) return fail(GLA_ERR_SCALE, "Volume pair must share a common scale grid.");
        d.src_div[a] = st > 0 ? st : 1;   // target cells coarser: source partitions
        d.trg_div[a] = st > 0 ? 1 : ts;
        if (trg->cel[a] % d.trg_div[a] || sv.cel[a] % d.src_div[a]) {
            if (trg->cel[a] < d.trg_div[a] || sv.cel[a] < d.src_div[a])
                return fail(GLA_ERR_SCALE, "Volume sizes are incompatible---volume smaller than cell.");
            return fail(GLA_ERR_SIZE, "a volume with finer cells must divide into whole partitions: its cell "
                                      "count on every axis must be a multiple of the scale ratio");
        }
    }
    const int64_t np = d.src_div[0] * d.src_div[1] * d.src_div[2] * d.trg_div[0] * d.trg_div[1] * d.trg_div[2];
    for (int a = 0; a < 3; ++a) {
        d.trg_cel[a] = trg->cel[a] / d.trg_div[a];
        d.src_cel[a] = sv.cel[a] / d.src_div[a];
        if ((d.trg_cel[a] + d.src_cel[a]) % 2)
            return fail(GLA_ERR_SIZE, np > 1 ? "the sum of source and target partition cells must be even on every "
                                               "axis"
                                             : "the sum of source and target cells must be even on every axis "
                                               "(regenerate the volumes with an even cell count first)");
        d.brn[a] = (d.trg_cel[a] + d.src_cel[a]) / 2;
        // an embedded axis stores its full branches
        const int b = (int)d.brn[a];
        const bool keep = mode == DIMS_REF || (mode == DIMS_EXACT ? gla::len_supported(b) : gla::len_fast(b));
        d.dev_brn[a] = keep ? d.brn[a] : gla::len_next_supported(b);
        if (d.dev_brn[a] == 0)
            return fail(GLA_ERR_UNSUPPORTED, "branch length " + std::to_string(d.brn[a]) + " exceeds the compiled transforms");
        const bool same_org = trg->org_num[a] * sv.org_den[a] == sv.org_num[a] * trg->org_den[a];
        d.sym[a] = (d.trg_cel[a] == d.src_cel[a] && same_org && d.dev_brn[a] == d.brn[a] && np == 1) ? 1 : 0;
        if (d.sym[a]) {
            int64_t u = (d.trg_cel[a] + 1) / 2 + (d.trg_cel[a] % 2 == 0);
            d.fur_ext[a] = u < 2 ? 2 : u;
        } else {
            d.fur_ext[a] = d.dev_brn[a];
        }
    }
    *out = d;
    return GLA_OK;
}

// source and target partition counts
int64_t n_src_part(const gla_dims& d) { return d.src_div[0] * d.src_div[1] * d.src_div[2]; }
int64_t n_trg_part(const gla_dims& d) { return d.trg_div[0] * d.trg_div[1] * d.trg_div[2]; }

// complex elements of the Green data in the reference layout
int64_t fur_count(const gla_dims& d)
{
    return 8 * d.fur_ext[0] * d.fur_ext[1] * d.fur_ext[2] * 6 * n_src_part(d) * n_trg_part(d);
}

}  // namespace

extern "C" {

gla_status gla_opr_dims_for(const gla_vol* trg, const gla_vol* src, gla_dims* out)
{
    if (!trg || !out) return fail(GLA_ERR_ARG, "null argument");
    gla_vol t2, s2;
    even_pair(trg, src, t2, s2);
    // the layout the caller hands in, with the lengths the GPU will use
    gla_dims ref, dev;
    gla_status s = dims_impl(&t2, src ? &s2 : nullptr, DIMS_REF, &ref);
    if (s) return s;
    if ((s = dims_impl(&t2, src ? &s2 : nullptr, DIMS_EXACT, &dev))) return s;
    for (int a = 0; a < 3; ++a) ref.dev_brn[a] = dev.dev_brn[a];
    *out = ref;
    return GLA_OK;
}

gla_status gla_vol_even_gen(const gla_vol* in, gla_vol* out)
{
    if (!in || !out) return fail(GLA_ERR_ARG, "null argument");
    gla_status s = check_vol(*in, "volume");
    if (s) return s;
    *out = even_gen(*in);
    return GLA_OK;
}

gla_status gla_opr_get_dims(gla_opr_t h, gla_dims* out)
{
    if (!h || !out) return fail(GLA_ERR_ARG, "null argument");
    *out = h->dims;
    return GLA_OK;
}

}  // extern "C"

namespace {

template <typename R>
gla_status upload(gla_opr_t h, const std::vector<std::complex<double>>& furd)
{
    using C = std::complex<R>;
    const gla_dims& d = h->dims;
    const int64_t U0 = d.fur_ext[0], U1 = d.fur_ext[1], U2 = d.fur_ext[2];
    const int64_t Ut = U0 * U1 * U2;
    const int64_t P = n_src_part(d) * n_trg_part(d);
    const double scale = 1.0 / (double)(8 * d.dev_brn[0] * d.dev_brn[1] * d.dev_brn[2]);
    // device layout, one block per partition pair p = s + Ps * t:
    // (((((p*8 + b)*U2 + j2)*U1 + j1)*U0 + j0)*6 + comp; the reference layout
    // is ((b*P + p)*6 + comp)*Ut + j
    std::vector<C> g((size_t)(P * 8 * Ut * 6));
    for (int64_t p = 0; p < P; ++p)
        for (int b = 0; b < 8; ++b)
            for (int c = 0; c < 6; ++c)
                for (int64_t j = 0; j < Ut; ++j) {
                    const std::complex<double> v = furd[(size_t)(((b * P + p) * 6 + c) * Ut + j)] * scale;
                    g[(size_t)(((p * 8 + b) * Ut + j) * 6 + c)] = C((R)v.real(), (R)v.imag());
                }
    // phase and twiddle tables per axis
    int64_t ntab = 0;
    for (int a = 0; a < 3; ++a) ntab += 2 * d.dev_brn[a];
    std::vector<C> tb((size_t)ntab);
    int64_t off[3][2];
    int64_t p = 0;
    for (int a = 0; a < 3; ++a) {
        const int64_t B = d.dev_brn[a];
        off[a][0] = p;
        for (int64_t n = 0; n < B; ++n) {
            const double t = -M_PI * (double)n / (double)B;
            tb[(size_t)(p++)] = C((R)std::cos(t), (R)std::sin(t));
        }
        off[a][1] = p;
        for (int64_t n = 0; n < B; ++n) {
            const double t = -2.0 * M_PI * (double)n / (double)B;
            tb[(size_t)(p++)] = C((R)std::cos(t), (R)std::sin(t));
        }
    }
    for (int a = 0; a < 3; ++a) {
        gla::AxisMap& m = h->tab.am[a];
        m.B = (int)d.dev_brn[a];
        m.U = (int)d.fur_ext[a];
        m.sym = (d.sym[a] && d.fur_ext[a] != d.brn[a]) ? 1 : 0;
        const int T = (int)d.trg_cel[a];
        m.fe = (T + 1) / 2 + (T % 2 == 0 ? 1 : 0);
        m.fo = (T + 1) / 2;
    }
    const int Bv[3] = {(int)d.dev_brn[0], (int)d.dev_brn[1], (int)d.dev_brn[2]};
    const int Lv[3] = {2 * Bv[0], 2 * Bv[1], 2 * Bv[2]};
    const std::vector<int> order = gla::zpass_order(Bv, Lv, h->tab.am, gla::zpass_tile_width((int)sizeof(C), Bv[2], Lv[0], Lv[1]));
    auto dev = std::make_shared<DevData>();
    dev->device = h->opts.device;
    GLA_HIP(hipMalloc(&dev->gf, g.size() * sizeof(C)));
    GLA_HIP(hipMalloc(&dev->tab, tb.size() * sizeof(C)));
    GLA_HIP(hipMalloc(&dev->order, order.size() * sizeof(int)));
    GLA_HIP(hipMemcpy(dev->gf, g.data(), g.size() * sizeof(C), hipMemcpyHostToDevice));
    GLA_HIP(hipMemcpy(dev->tab, tb.data(), tb.size() * sizeof(C), hipMemcpyHostToDevice));
    GLA_HIP(hipMemcpy(dev->order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    h->dev = dev;
    for (int a = 0; a < 3; ++a) {
        h->tab.phz[a] = (const C*)dev->tab + off[a][0];
        h->tab.tw[a] = (const C*)dev->tab + off[a][1];
    }
    h->tab.gf = dev->gf;
    h->tab.zorder = (const int*)dev->order;
    h->tab.zper_xcd = (int)(order.size() / 8);
    return GLA_OK;
}

// launch geometry of partition pair (0, 0); other pairs change only the
// partition offsets (pair_geom)
void fill_geom(gla_opr_t h)
{
    gla::Geom& g = h->geom;
    const gla_dims& d = h->dims;
    int fs[3], ft[3];
    for (int a = 0; a < 3; ++a) {
        g.S[a] = (int)d.src_cel[a];
        g.T[a] = (int)d.trg_cel[a];
        g.B[a] = (int)d.dev_brn[a];
        g.L[a] = 2 * g.B[a];
        fs[a] = (int)(d.src_cel[a] * d.src_div[a]);
        ft[a] = (int)(d.trg_cel[a] * d.trg_div[a]);
    }
    g.Y1 = std::max(g.S[1], g.T[1]);
    g.Z = std::max(g.S[2], g.T[2]);
    g.xs = (int64_t)fs[0] * fs[1] * fs[2] * 3;
    g.ys = (int64_t)ft[0] * ft[1] * ft[2] * 3;
    g.w1s = (int64_t)g.L[0] * g.Y1 * g.Z * 3;
    g.w2s = (int64_t)g.L[0] * g.L[1] * g.Z * 3;
    g.xp = gla::whole_part(fs);
    g.yp = gla::whole_part(ft);
    for (int a = 0; a < 3; ++a) {
        g.xp.d[a] = (int)d.src_div[a];
        g.yp.d[a] = (int)d.trg_div[a];
    }
}

// partition offsets of partition `p` on a grid of divisors div (x fastest,
// CartesianIndices order of GlaExtInf's srcPar / trgPar)
void part_offset(int64_t p, const int64_t div[3], int o[3])
{
    o[0] = (int)(p % div[0]);
    o[1] = (int)((p / div[0]) % div[1]);
    o[2] = (int)(p / (div[0] * div[1]));
}

size_t elem_bytes(int dtype) { return dtype == GLA_C64 ? 8 : 16; }

// The whole matvec on `st` with work planes w1 / w2.  A partitioned operator
// runs the five passes once per (source, target) partition pair: pass 1
// gathers source partition s (genPrt), pass 5 writes or accumulates into
// target partition t (mrgPrt); the first source partition of a target
// partition applies (alpha, beta), the later ones add with beta = 1, which
// is the leaf sum of egoBrnHst! (src/glaAct.jl:158-164) carried through the
// inverse transforms (they are linear).
hipError_t matvec_all(const gla_opr_s& h, const void* x, void* y, void* w1, void* w2, int nrhs, double ar,
                      double ai, double br, double bi, int axpby, const gla::Epi& epi, hipStream_t st)
{
    const bool c64 = h.dims.dtype == GLA_C64;
    auto one = [&](const gla::Geom& g, const gla::DevTab& tb, double b_r, double b_i, int ax) {
        return c64 ? gla::run_matvec<float>(g, tb, x, y, w1, w2, nrhs, ar, ai, b_r, b_i, ax, epi, st)
                   : gla::run_matvec<double>(g, tb, x, y, w1, w2, nrhs, ar, ai, b_r, b_i, ax, epi, st);
    };
    const int64_t Ps = n_src_part(h.dims), Pt = n_trg_part(h.dims);
    if (Ps * Pt == 1) return one(h.geom, h.tab, br, bi, axpby);
    if (axpby == 2) return hipErrorInvalidValue;   // self operators only, never partitioned
    if (axpby == 0) {
        ar = 1; ai = 0; br = 0; bi = 0;
    }
    const int64_t blk = 8 * h.dims.fur_ext[0] * h.dims.fur_ext[1] * h.dims.fur_ext[2] * 6 *
                        (int64_t)elem_bytes(h.dims.dtype);
    for (int64_t t = 0; t < Pt; ++t)
        for (int64_t s = 0; s < Ps; ++s) {
            gla::Geom g = h.geom;
            part_offset(s, h.dims.src_div, g.xp.o);
            part_offset(t, h.dims.trg_div, g.yp.o);
            gla::DevTab tb = h.tab;
            tb.gf = (const char*)h.tab.gf + (s + Ps * t) * blk;
            const hipError_t e = one(g, tb, s == 0 ? br : 1.0, s == 0 ? bi : 0.0, 1);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

gla_status ensure_ws(gla_opr_t h, int64_t nrhs)
{
    if (h->ws_nrhs >= nrhs) return GLA_OK;
    if (h->batches > 0)
        return fail(GLA_ERR_ARG, "the operator's work buffers are recorded in a batch; "
                                 "destroy the batch before applying it to more right-hand sides");
    const size_t eb = elem_bytes(h->dims.dtype);
    if (h->w1) { (void)hipFree(h->w1); h->w1 = nullptr; }
    if (h->w2) { (void)hipFree(h->w2); h->w2 = nullptr; }
    h->ws_nrhs = 0;
    GLA_HIP(hipMalloc(&h->w1, (size_t)(h->geom.w1s * nrhs) * eb));
    GLA_HIP(hipMalloc(&h->w2, (size_t)(h->geom.w2s * nrhs) * eb));
    h->ws_nrhs = nrhs;
    return GLA_OK;
}

gla_status run(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs, double ar,
               double ai, double br, double bi, int axpby, void* stream, const gla::Epi& epi = gla::Epi())
{
    if (!h) return fail(GLA_ERR_ARG, "null operator");
    if (nrhs < 1 || nrhs > 65535) return fail(GLA_ERR_ARG, "nrhs must be in [1, 65535]");
    if (!x || !y) return fail(GLA_ERR_ARG, "null vector");
    if (nx != h->geom.xs * nrhs)
        return fail(GLA_ERR_SIZE, "Size of input vector does not match size of source volume. Required data "
                                  "layout is (srcCelX, srcCelY, srcCelZ, 3).");
    if (ny != h->geom.ys * nrhs) return fail(GLA_ERR_SIZE, "Size of output vector does not match the target volume.");
    std::lock_guard<std::mutex> lk(h->mu);
    DevGuard g(h->opts.device);
    if (h->ws_nrhs < nrhs) {
        // growing the work buffers frees and allocates, which a stream under
        // graph capture must not see
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (stream && hipStreamIsCapturing((hipStream_t)stream, &cs) == hipSuccess &&
            cs != hipStreamCaptureStatusNone)
            return fail(GLA_ERR_ARG, "work buffers for " + std::to_string(nrhs) +
                                         " right-hand sides must be reserved (gla_opr_reserve) before capture");
    }
    gla_status s = ensure_ws(h, nrhs);
    if (s) return s;
    const hipError_t e = matvec_all(*h, x, y, h->w1, h->w2, (int)nrhs, ar, ai, br, bi, axpby, epi,
                                    (hipStream_t)stream);
    if (e != hipSuccess) return fail(GLA_ERR_HIP, std::string("matvec launch: ") + hipGetErrorString(e));
    return GLA_OK;
}

template <typename T>
bool finite_all(const T* p, int64_t n)
{
    for (int64_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

}  // namespace

namespace {
gla_status create_common(const gla_opts* opts, const gla_vol* trg, const gla_vol* src, gla_dims d, const void* fur,
                         int32_t fur_dtype, int64_t fur_len, gla_opr_t* out);
}

extern "C" {

gla_status gla_opr_create_from_fur(const gla_opts* opts, const gla_vol* trg, const gla_vol* src,
                                   const void* fur, int32_t fur_dtype, int64_t fur_len, gla_opr_t* out)
{
    if (!opts || !trg || !fur || !out) return fail(GLA_ERR_ARG, "null argument");
    *out = nullptr;
    if (opts->dtype != GLA_C64 && opts->dtype != GLA_C128)
        return fail(GLA_ERR_ARG, "setTyp must be a subtype of Complex (ComplexF32 or ComplexF64)");
    if (fur_dtype != GLA_C64 && fur_dtype != GLA_C128) return fail(GLA_ERR_ARG, "fur dtype must be c64 or c128");
    gla_status s = gla_init(opts->device);
    if (s) return s;
    gla_vol t2, s2;
    even_pair(trg, src, t2, s2);
    const gla_vol* sp = src ? &s2 : nullptr;
    gla_dims ref, d;
    if ((s = dims_impl(&t2, sp, DIMS_REF, &ref))) return s;
    if ((s = dims_impl(&t2, sp, DIMS_EXACT, &d))) return s;
    bool same = true;
    for (int a = 0; a < 3; ++a) same = same && d.dev_brn[a] == ref.brn[a];
    if (same) return create_common(opts, &t2, sp, d, fur, fur_dtype, fur_len, out);
    // a branch length without a compiled transform: the data is re-expressed
    // at the next fast length (gla::refur) in FP64
    const int64_t Ut = ref.fur_ext[0] * ref.fur_ext[1] * ref.fur_ext[2];
    if (fur_len != fur_count(ref))
        return fail(GLA_ERR_SIZE, "Fourier Green data has " + std::to_string(fur_len) + " elements, expected 8*" +
                                      std::to_string(Ut) + "*6*" + std::to_string(n_src_part(ref) * n_trg_part(ref)));
    const bool ok = fur_dtype == GLA_C64 ? finite_all((const float*)fur, fur_len)
                                         : finite_all((const double*)fur, fur_len);
    if (!ok) return fail(GLA_ERR_NONFINITE, "Fourier information contains non-numeric values.");
    std::vector<std::complex<double>> furd((size_t)fur_len), re;
    for (int64_t i = 0; i < fur_len; ++i)
        furd[(size_t)i] = fur_dtype == GLA_C64
                              ? std::complex<double>(((const float*)fur)[2 * i], ((const float*)fur)[2 * i + 1])
                              : std::complex<double>(((const double*)fur)[2 * i], ((const double*)fur)[2 * i + 1]);
    std::string err;
    {
        DevGuard g(opts->device);
        if ((s = gla::refur(ref, d, furd, re, err))) return fail(s, err);
    }
    return create_common(opts, &t2, sp, d, re.data(), GLA_C128, (int64_t)re.size(), out);
}

}  // extern "C"

namespace {

gla_status create_common(const gla_opts* opts, const gla_vol* trg, const gla_vol* src, gla_dims d, const void* fur,
                         int32_t fur_dtype, int64_t fur_len, gla_opr_t* out)
{
    gla_status s;
    const int64_t Ut = d.fur_ext[0] * d.fur_ext[1] * d.fur_ext[2];
    if (fur_len != fur_count(d))
        return fail(GLA_ERR_SIZE, "Fourier Green data has " + std::to_string(fur_len) + " elements, expected 8*" +
                                      std::to_string(Ut) + "*6*" + std::to_string(n_src_part(d) * n_trg_part(d)));
    const bool ok = fur_dtype == GLA_C64 ? finite_all((const float*)fur, fur_len)
                                         : finite_all((const double*)fur, fur_len);
    if (!ok) return fail(GLA_ERR_NONFINITE, "Fourier information contains non-numeric values.");
    d.dtype = opts->dtype;
    d.adj_mod = opts->adj_mod;
    std::vector<std::complex<double>> furd((size_t)fur_len);
    for (int64_t i = 0; i < fur_len; ++i)
        furd[(size_t)i] = fur_dtype == GLA_C64
                              ? std::complex<double>(((const float*)fur)[2 * i], ((const float*)fur)[2 * i + 1])
                              : std::complex<double>(((const double*)fur)[2 * i], ((const double*)fur)[2 * i + 1]);
    // the Green data in the operator's element type first, as the
    // reference's setTyp cast stores it (src/glaMemSup.jl:148): the device
    // copy is then the same whether the data came from the setup, from
    // egoFur or from a file
    if (opts->dtype == GLA_C64)
        for (auto& v : furd) v = std::complex<double>((float)v.real(), (float)v.imag());
    auto h = std::make_unique<gla_opr_s>();
    h->opts = *opts;
    h->trg = *trg;
    h->src = src ? *src : *trg;
    h->dims = d;
    fill_geom(h.get());
    {
        DevGuard g(opts->device);
        s = opts->dtype == GLA_C64 ? upload<float>(h.get(), furd) : upload<double>(h.get(), furd);
        if (s) return s;
    }
    // host copy in the operator's element type, rounded as the reference's
    // setTyp cast rounds it (src/glaMemSup.jl:148)
    auto fh = std::make_shared<std::vector<char>>((size_t)fur_len * elem_bytes(opts->dtype));
    for (int64_t i = 0; i < fur_len; ++i) {
        if (opts->dtype == GLA_C64) {
            ((float*)fh->data())[2 * i] = (float)furd[(size_t)i].real();
            ((float*)fh->data())[2 * i + 1] = (float)furd[(size_t)i].imag();
        } else {
            ((double*)fh->data())[2 * i] = furd[(size_t)i].real();
            ((double*)fh->data())[2 * i + 1] = furd[(size_t)i].imag();
        }
    }
    h->fur_host = fh;
    *out = h.release();
    return GLA_OK;
}

}  // namespace

extern "C" {

gla_status gla_opr_create(const gla_opts* opts, const gla_vol* trg, const gla_vol* src, gla_opr_t* out)
{
    if (!opts || !trg || !out) return fail(GLA_ERR_ARG, "null argument");
    *out = nullptr;
    if (opts->dtype != GLA_C64 && opts->dtype != GLA_C128)
        return fail(GLA_ERR_ARG, "setTyp must be a subtype of Complex (ComplexF32 or ComplexF64)");
    gla_status s = gla_init(opts->device);
    if (s) return s;
    gla_vol t2, s2;
    even_pair(trg, src, t2, s2);
    const gla_vol* sp = src ? &s2 : nullptr;
    gla_dims d;
    if ((s = dims_impl(&t2, sp, GLA_CREATE_DIMS, &d))) return s;
    std::vector<std::complex<double>> fur;
    std::string err;
    s = gla::setup_fur(*opts, t2, sp ? s2 : t2, d, fur, err);
    if (s) return fail(s, err);
    return create_common(opts, &t2, sp, d, fur.data(), GLA_C128, (int64_t)fur.size(), out);
}

gla_status gla_opr_export_fur(gla_opr_t h, void* buf, int64_t len)
{
    if (!h || !buf) return fail(GLA_ERR_ARG, "null argument");
    const int64_t n = (int64_t)(h->fur_host->size() / elem_bytes(h->dims.dtype));
    if (len < n) return fail(GLA_ERR_SIZE, "export buffer holds " + std::to_string(len) + " elements, need " +
                                              std::to_string(n));
    std::memcpy(buf, h->fur_host->data(), h->fur_host->size());
    return GLA_OK;
}

gla_status gla_opr_get_vols(gla_opr_t h, gla_vol* trg, gla_vol* src)
{
    if (!h || !trg || !src) return fail(GLA_ERR_ARG, "null argument");
    *trg = h->trg;
    *src = h->src;
    return GLA_OK;
}

gla_status gla_opr_get_opts(gla_opr_t h, gla_opts* opts)
{
    if (!h || !opts) return fail(GLA_ERR_ARG, "null argument");
    *opts = h->opts;
    return GLA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Green-data file: a 512-byte little-endian header, then the eight branches
// in the export layout (the operator's element type, re/im interleaved).
// The header carries everything that fixes the layout (volumes, branch and
// device lengths, stored extents, symmetric axes) so a reader can check that
// the payload belongs to the geometry it names.
// ---------------------------------------------------------------------------
namespace {

constexpr char kFurMagic[8] = {'G', 'I', 'L', 'A', 'F', 'U', 'R', '1'};
constexpr uint32_t kFurVersion = 1;
constexpr uint32_t kFurHeader = 512;

struct FurHeader {
    char magic[8];
    uint32_t version, header_bytes;
    int32_t dtype, int_ord, adj_mod, self_opr;
    double frq_re, frq_im;
    gla_vol trg, src;
    int64_t brn[3], dev_brn[3], fur_ext[3];
    int32_t sym[3], reserved;
    uint64_t n_elements, fnv1a;
};
static_assert(sizeof(FurHeader) <= kFurHeader, "Green-data header exceeds its 512 bytes");

uint64_t fnv1a64(const unsigned char* p, size_t n)
{
    uint64_t hsh = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        hsh ^= p[i];
        hsh *= 1099511628211ull;
    }
    return hsh;
}

bool little_endian_host()
{
    const uint32_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

}  // namespace

extern "C" {

gla_status gla_opr_save(gla_opr_t h, const char* path)
{
    if (!h || !path) return fail(GLA_ERR_ARG, "null argument");
    if (!little_endian_host()) return fail(GLA_ERR_UNSUPPORTED, "Green-data files are written on little-endian hosts");
    FurHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, kFurMagic, 8);
    hd.version = kFurVersion;
    hd.header_bytes = kFurHeader;
    hd.dtype = h->dims.dtype;
    hd.int_ord = h->opts.int_ord;
    hd.adj_mod = h->opts.adj_mod;
    hd.self_opr = h->dims.self_opr;
    hd.frq_re = h->opts.frq_phz_re;
    hd.frq_im = h->opts.frq_phz_im;
    hd.trg = h->trg;
    hd.src = h->src;
    for (int a = 0; a < 3; ++a) {
        hd.brn[a] = h->dims.brn[a];
        hd.dev_brn[a] = h->dims.dev_brn[a];
        hd.fur_ext[a] = h->dims.fur_ext[a];
        hd.sym[a] = h->dims.sym[a];
    }
    const std::vector<char>& pay = *h->fur_host;
    hd.n_elements = pay.size() / elem_bytes(h->dims.dtype);
    hd.fnv1a = fnv1a64((const unsigned char*)pay.data(), pay.size());
    std::vector<char> head(kFurHeader, 0);
    std::memcpy(head.data(), &hd, sizeof hd);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(GLA_ERR_ARG, std::string("cannot open ") + path + " for writing");
    const bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                    std::fwrite(pay.data(), 1, pay.size(), f) == pay.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(GLA_ERR_ARG, std::string("short write to ") + path);
    return GLA_OK;
}

gla_status gla_opr_load(const char* path, int32_t dtype, int32_t device, gla_opr_t* out)
{
    if (!path || !out) return fail(GLA_ERR_ARG, "null argument");
    *out = nullptr;
    if (!little_endian_host()) return fail(GLA_ERR_UNSUPPORTED, "Green-data files are read on little-endian hosts");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(GLA_ERR_ARG, std::string("cannot open ") + path);
    std::vector<char> head(kFurHeader);
    FurHeader hd;
    if (std::fread(head.data(), 1, head.size(), f) != head.size()) {
        std::fclose(f);
        return fail(GLA_ERR_SIZE, std::string(path) + ": shorter than a Green-data header");
    }
    std::memcpy(&hd, head.data(), sizeof hd);
    auto bad = [&](const std::string& why) {
        std::fclose(f);
        return fail(GLA_ERR_ARG, std::string(path) + ": " + why);
    };
    if (std::memcmp(hd.magic, kFurMagic, 8) != 0) return bad("not a Green-data file");
    if (hd.version != kFurVersion || hd.header_bytes != kFurHeader) return bad("unsupported Green-data version");
    if (hd.dtype != GLA_C64 && hd.dtype != GLA_C128) return bad("unknown element type");
    gla_dims d;
    const gla_vol* sp = hd.self_opr ? nullptr : &hd.src;
    // the file holds either the layout of Fourier data handed in or that of
    // an operator built from geometry
    bool match = false;
    for (int exact = 0; exact < 2 && !match; ++exact) {
        if (dims_impl(&hd.trg, sp, exact != 0 ? DIMS_EXACT : DIMS_FAST, &d) != GLA_OK)
            return bad("volumes in the header are inconsistent");
        match = true;
        for (int a = 0; a < 3; ++a)
            if (d.brn[a] != hd.brn[a] || d.dev_brn[a] != hd.dev_brn[a] || d.fur_ext[a] != hd.fur_ext[a] ||
                d.sym[a] != hd.sym[a])
                match = false;
    }
    if (!match) return bad("branch layout in the header does not match its volumes");
    const uint64_t want = (uint64_t)fur_count(d);
    if (hd.n_elements != want) return bad("element count does not match the branch layout");
    const size_t eb = hd.dtype == GLA_C64 ? 8 : 16;
    std::vector<char> pay((size_t)want * eb);
    const size_t got = std::fread(pay.data(), 1, pay.size(), f);
    char extra;
    const bool trailing = std::fread(&extra, 1, 1, f) == 1;
    std::fclose(f);
    if (got != pay.size()) return fail(GLA_ERR_SIZE, std::string(path) + ": truncated Green data");
    if (trailing) return fail(GLA_ERR_SIZE, std::string(path) + ": trailing bytes after the Green data");
    if (fnv1a64((const unsigned char*)pay.data(), pay.size()) != hd.fnv1a)
        return fail(GLA_ERR_ARG, std::string(path) + ": checksum mismatch");
    gla_opts o;
    o.frq_phz_re = hd.frq_re;
    o.frq_phz_im = hd.frq_im;
    o.int_ord = hd.int_ord;
    o.adj_mod = hd.adj_mod;
    o.dtype = dtype == 0 ? hd.dtype : dtype;
    o.device = device;
    if (o.dtype != GLA_C64 && o.dtype != GLA_C128) return fail(GLA_ERR_TYPE, "dtype must be 0, GLA_C64 or GLA_C128");
    gla_status s = gla_init(device);
    if (s) return s;
    return create_common(&o, &hd.trg, sp, d, pay.data(), hd.dtype, (int64_t)want, out);
}

gla_status gla_opr_adjoint(gla_opr_t h, gla_opr_t* out)
{
    if (!h || !out) return fail(GLA_ERR_ARG, "null argument");
    auto a = std::make_unique<gla_opr_s>();
    a->opts = h->opts;
    a->opts.adj_mod = !h->opts.adj_mod;
    // target and source volumes swap; the partition info and the Green data
    // are kept as they are (src/glaLinAlg.jl:16-29)
    a->trg = h->src;
    a->src = h->trg;
    a->dims = h->dims;
    a->dims.adj_mod = a->opts.adj_mod;
    a->geom = h->geom;
    a->tab = h->tab;
    a->dev = h->dev;
    a->fur_host = h->fur_host;
    *out = a.release();
    return GLA_OK;
}

gla_status gla_opr_destroy(gla_opr_t h)
{
    delete h;
    return GLA_OK;
}

gla_status gla_opr_apply(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs, void* stream)
{
    return run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 0, stream);
}

gla_status gla_opr_apply_noalloc(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                                 void* stream)
{
    if (!h) return fail(GLA_ERR_ARG, "null operator");
    if (nrhs < 1 || nrhs > 65535) return fail(GLA_ERR_ARG, "nrhs must be in [1, 65535]");
    if (!x || !y) return fail(GLA_ERR_ARG, "null vector");
    if (nx != h->geom.xs * nrhs)
        return fail(GLA_ERR_SIZE, "Size of input vector does not match size of source volume. Required data "
                                  "layout is (srcCelX, srcCelY, srcCelZ, 3).");
    if (ny != h->geom.ys * nrhs) return fail(GLA_ERR_SIZE, "Size of output vector does not match the target volume.");
    const size_t eb = elem_bytes(h->dims.dtype);
    // W1 then W2, each start on a 256-byte boundary
    const size_t w1b = ((size_t)(h->geom.w1s * nrhs) * eb + 255) & ~(size_t)255;
    const size_t need = w1b + (size_t)(h->geom.w2s * nrhs) * eb;
    if (need > kScratchBytes)
        return fail(GLA_ERR_SIZE, "the allocation-free matvec needs " + std::to_string(need) +
                                      " bytes of work memory, more than its fixed " + std::to_string(kScratchBytes) +
                                      "-byte scratch (use gla_opr_apply)");
    void* base = nullptr;
    gla_status s = scratch_addr(h->opts.device, &base);
    if (s) return s;
    std::lock_guard<std::mutex> lk(h->mu);
    DevGuard g(h->opts.device);
    void* w1 = base;
    void* w2 = (unsigned char*)base + w1b;
    const hipError_t e = matvec_all(*h, x, y, w1, w2, (int)nrhs, 1, 0, 0, 0, 0, gla::Epi(), (hipStream_t)stream);
    if (e != hipSuccess) return fail(GLA_ERR_HIP, std::string("matvec launch: ") + hipGetErrorString(e));
    return GLA_OK;
}

gla_status gla_opr_apply_axpby(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                               double alpha_re, double alpha_im, double beta_re, double beta_im, void* stream)
{
    return run(h, x, nx, y, ny, nrhs, alpha_re, alpha_im, beta_re, beta_im, 1, stream);
}

gla_status gla_ls_apply(gla_opr_t h, const void* chi, int64_t nchi, int64_t chi_stride, const void* x, int64_t nx,
                        void* y, int64_t ny, int64_t nrhs, void* stream)
{
    if (!h || !chi) return fail(GLA_ERR_ARG, "null argument");
    if (n_src_part(h->dims) * n_trg_part(h->dims) != 1)
        return fail(GLA_ERR_SIZE, "the Lippmann-Schwinger form needs equal source and target volumes");
    for (int a = 0; a < 3; ++a)
        if (h->geom.S[a] != h->geom.T[a])
            return fail(GLA_ERR_SIZE, "the Lippmann-Schwinger form needs equal source and target volumes");
    const int64_t cells = (int64_t)h->geom.T[0] * h->geom.T[1] * h->geom.T[2];
    if (chi_stride != 0 && chi_stride < cells) return fail(GLA_ERR_ARG, "chi_stride must be 0 or >= the cell count");
    const int64_t need = chi_stride == 0 ? cells : chi_stride * (nrhs - 1) + cells;
    if (nchi < need)
        return fail(GLA_ERR_SIZE, "Green's operator and medium must have the same size.");
    if (x == y) return fail(GLA_ERR_ARG, "x and y must not alias");
    gla::Epi epi;
    epi.x = x;
    epi.xs = h->geom.xs;
    epi.chi = chi;
    epi.chis = chi_stride;
    return run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 2, stream, epi);
}

gla_status gla_opr_apply_host(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs)
{
    if (!h) return fail(GLA_ERR_ARG, "null operator");
    if (nrhs < 1 || nx != h->geom.xs * nrhs || ny != h->geom.ys * nrhs || !x || !y)
        return run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 0, nullptr);   // raises the argument or size error
    std::lock_guard<std::mutex> lk(h->hmu);
    DevGuard g(h->opts.device);
    const size_t eb = elem_bytes(h->dims.dtype);
    auto grow = [&](void*& p, int64_t& have, int64_t want) -> hipError_t {
        if (have >= want) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        have = 0;
        const hipError_t e = hipMalloc(&p, (size_t)want * eb);
        if (e == hipSuccess) have = want;
        return e;
    };
    hipError_t e = grow(h->hx, h->hx_n, nx);
    if (e == hipSuccess) e = grow(h->hy, h->hy_n, ny);
    if (e != hipSuccess) return fail(GLA_ERR_HIP, hipGetErrorString(e));
    gla_status s = GLA_OK;
    e = hipMemcpy(h->hx, x, (size_t)nx * eb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        s = run(h, h->hx, nx, h->hy, ny, nrhs, 1, 0, 0, 0, 0, nullptr);
        if (s == GLA_OK) e = hipMemcpy(y, h->hy, (size_t)ny * eb, hipMemcpyDeviceToHost);
    }
    if (s) return s;
    if (e != hipSuccess) return fail(GLA_ERR_HIP, hipGetErrorString(e));
    return GLA_OK;
}

}  // extern "C"
namespace {
gla_status profile_locked(gla_opr_t h, const void* x, void* y, int64_t nrhs, void* stream, int32_t reps, double* ms,
                          int64_t* bytes);
void pass_bytes(gla_opr_t h, int64_t nrhs, int64_t* bytes);
}
extern "C" {

gla_status gla_opr_profile(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs, void* stream,
                           int32_t reps, double* ms, int64_t* bytes)
{
    if (!h || !ms || !bytes || reps < 1) return fail(GLA_ERR_ARG, "bad profile arguments");
    if (n_src_part(h->dims) * n_trg_part(h->dims) != 1)
        return fail(GLA_ERR_UNSUPPORTED, "per-pass profiles cover operators with one partition pair");
    gla_status s = run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 0, stream);   // validates, allocates, warms up
    if (s) return s;
    s = profile_locked(h, x, y, nrhs, stream, reps, ms, bytes);
    if (s) return s;
    // leave y holding a complete matvec again (pass 2 ran in place reps times)
    return run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 0, stream);
}

gla_status gla_opr_profile_seq(gla_opr_t h, const void* x, int64_t nx, void* y, int64_t ny, int64_t nrhs,
                               void* stream, int32_t reps, double* ms, int64_t* bytes, double* total_ms)
{
    if (!h || !ms || !bytes || !total_ms || reps < 1 || reps > 100000) return fail(GLA_ERR_ARG, "bad profile arguments");
    if (n_src_part(h->dims) * n_trg_part(h->dims) != 1)
        return fail(GLA_ERR_UNSUPPORTED, "per-pass profiles cover operators with one partition pair");
    gla_status s = run(h, x, nx, y, ny, nrhs, 1, 0, 0, 0, 0, stream);   // validates, allocates, warms up
    if (s) return s;
    std::lock_guard<std::mutex> lk(h->mu);
    DevGuard g(h->opts.device);
    pass_bytes(h, nrhs, bytes);   // algorithmic bytes per pass, as gla_opr_profile
    // reps matvecs exactly as gla_opr_apply issues them, with an event at
    // every pass boundary: pass p of matvec r lasts ev[r][p] -> ev[r][p+1]
    const int ne = gla::NPASS + 1;
    std::vector<hipEvent_t> ev((size_t)reps * ne, nullptr);
    hipError_t e = hipSuccess;
    for (auto& v : ev)
        if (e == hipSuccess) e = hipEventCreate(&v);
    hipStream_t st = (hipStream_t)stream;
    const gla::Geom& G = h->geom;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        e = hipEventRecord(ev[(size_t)r * ne], st);
        for (int p = 0; p < gla::NPASS && e == hipSuccess; ++p) {
            e = h->dims.dtype == GLA_C64
                    ? gla::run_pass<float>(p, G, h->tab, x, y, h->w1, h->w2, (int)nrhs, 1, 0, 0, 0, 0, gla::Epi(), st)
                    : gla::run_pass<double>(p, G, h->tab, x, y, h->w1, h->w2, (int)nrhs, 1, 0, 0, 0, 0, gla::Epi(),
                                            st);
            if (e == hipSuccess) e = hipEventRecord(ev[(size_t)r * ne + p + 1], st);
        }
    }
    if (e == hipSuccess) e = hipEventSynchronize(ev.back());
    double acc[gla::NPASS] = {0, 0, 0, 0, 0};
    for (int r = 0; r < reps && e == hipSuccess; ++r)
        for (int p = 0; p < gla::NPASS && e == hipSuccess; ++p) {
            float t = 0;
            e = hipEventElapsedTime(&t, ev[(size_t)r * ne + p], ev[(size_t)r * ne + p + 1]);
            acc[p] += t;
        }
    float tot = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev.front(), ev.back());
    for (auto& v : ev)
        if (v) (void)hipEventDestroy(v);
    if (e != hipSuccess) return fail(GLA_ERR_HIP, std::string("profile: ") + hipGetErrorString(e));
    for (int p = 0; p < gla::NPASS; ++p) ms[p] = acc[p] / reps;
    *total_ms = (double)tot / reps;
    return GLA_OK;
}

}  // extern "C"

namespace {

// algorithmic HBM bytes of each pass of one matvec over nrhs right-hand sides
void pass_bytes(gla_opr_t h, int64_t nrhs, int64_t* bytes)
{
    const gla::Geom& G = h->geom;
    const int64_t eb = (int64_t)elem_bytes(h->dims.dtype);
    const int64_t S = (int64_t)G.S[0] * G.S[1] * G.S[2], T = (int64_t)G.T[0] * G.T[1] * G.T[2];
    const int64_t a1 = (int64_t)G.L[0] * G.S[1] * G.S[2], a2 = (int64_t)G.L[0] * G.L[1] * G.S[2];
    const int64_t a3 = (int64_t)G.L[0] * G.L[1] * G.T[2], a4 = (int64_t)G.L[0] * G.T[1] * G.T[2];
    const int64_t gsz = 8 * h->dims.fur_ext[0] * h->dims.fur_ext[1] * h->dims.fur_ext[2] * 6;
    const int64_t f = 3 * nrhs * eb;
    bytes[0] = f * (S + a1);
    bytes[1] = f * (a1 + a2);
    bytes[2] = f * (a2 + a3) + gsz * eb;
    bytes[3] = f * (a3 + a4);
    bytes[4] = f * (a4 + T);
}

gla_status profile_locked(gla_opr_t h, const void* x, void* y, int64_t nrhs, void* stream, int32_t reps, double* ms,
                          int64_t* bytes)
{
    std::lock_guard<std::mutex> lk(h->mu);
    DevGuard g(h->opts.device);
    const gla::Geom& G = h->geom;
    pass_bytes(h, nrhs, bytes);
    hipEvent_t e0, e1;
    GLA_HIP(hipEventCreate(&e0));
    GLA_HIP(hipEventCreate(&e1));
    hipStream_t st = (hipStream_t)stream;
    for (int p = 0; p < gla::NPASS; ++p) {
        hipError_t e = hipSuccess;
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && e == hipSuccess; ++r)
            e = h->dims.dtype == GLA_C64
                    ? gla::run_pass<float>(p, G, h->tab, x, y, h->w1, h->w2, (int)nrhs, 1, 0, 0, 0, 0, gla::Epi(), st)
                    : gla::run_pass<double>(p, G, h->tab, x, y, h->w1, h->w2, (int)nrhs, 1, 0, 0, 0, 0, gla::Epi(),
                                            st);
        (void)hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return fail(GLA_ERR_HIP, std::string("profile: ") + hipGetErrorString(e));
        }
        ms[p] = (double)t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return GLA_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

// A fixed batch of matvecs recorded into one HIP graph: the operators are
// dealt over `lanes` parallel branches (largest first, each to the lane with
// the least work so far, an operator listed twice always to the same lane
// because both share its work buffers), so independent operators too small
// to fill the GPU run side by side; a replay is one graph launch.
struct gla_batch_s {
    int device = 0;
    std::vector<gla_opr_t> ops;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~gla_batch_s()
    {
        DevGuard g(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (gla_opr_t h : ops) {
            std::lock_guard<std::mutex> lk(h->mu);
            --h->batches;
        }
    }
};

extern "C" {

gla_status gla_batch_create(const gla_opr_t* ops, const void* const* x, void* const* y, const int64_t* nrhs,
                            int32_t n, int32_t lanes, gla_batch_t* out)
{
    if (!ops || !x || !y || !out || n < 1) return fail(GLA_ERR_ARG, "empty or null batch");
    *out = nullptr;
    if (lanes < 1) lanes = 1;
    if (lanes > n) lanes = n;
    const int dev = ops[0] ? ops[0]->opts.device : 0;
    for (int i = 0; i < n; ++i) {
        if (!ops[i] || !x[i] || !y[i]) return fail(GLA_ERR_ARG, "null operator or vector in the batch");
        if (ops[i]->opts.device != dev) return fail(GLA_ERR_ARG, "a batch runs on one device");
        const int64_t r = nrhs ? nrhs[i] : 1;
        if (r < 1 || r > 65535) return fail(GLA_ERR_ARG, "nrhs must be in [1, 65535]");
    }
    DevGuard g(dev);
    // work buffers first: no allocation may happen while recording
    for (int i = 0; i < n; ++i) {
        std::lock_guard<std::mutex> lk(ops[i]->mu);
        gla_status s = ensure_ws(ops[i], nrhs ? nrhs[i] : 1);
        if (s) return s;
    }
    // lane assignment by algorithmic bytes
    std::vector<int> lane((size_t)n, -1), idx((size_t)n);
    std::vector<double> cost((size_t)n), load((size_t)lanes, 0.0);
    for (int i = 0; i < n; ++i) {
        const gla::Geom& G = ops[i]->geom;
        cost[(size_t)i] = (double)(G.w1s + G.w2s) * (double)(nrhs ? nrhs[i] : 1);
        idx[(size_t)i] = i;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[(size_t)a] > cost[(size_t)b]; });
    for (int i : idx) {
        int l = -1;
        for (int j = 0; j < n; ++j)
            if (j != i && ops[j] == ops[i] && lane[(size_t)j] >= 0) l = lane[(size_t)j];
        if (l < 0) {
            l = 0;
            for (int k = 1; k < lanes; ++k)
                if (load[(size_t)k] < load[(size_t)l]) l = k;
        }
        lane[(size_t)i] = l;
        load[(size_t)l] += cost[(size_t)i];
    }
    auto b = std::make_unique<gla_batch_s>();
    b->device = dev;
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> ls((size_t)lanes, nullptr);
    std::vector<hipEvent_t> ev((size_t)lanes + 1, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    for (int k = 0; k < lanes && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&ls[(size_t)k], hipStreamNonBlocking);
    for (auto& v : ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&v, hipEventDisableTiming);
    bool capturing = false;
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        capturing = true;
        e = hipEventRecord(ev[(size_t)lanes], cap);   // fork
    }
    for (int k = 0; k < lanes && e == hipSuccess; ++k) e = hipStreamWaitEvent(ls[(size_t)k], ev[(size_t)lanes], 0);
    // record in list order within each lane
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        gla_opr_t h = ops[i];
        std::lock_guard<std::mutex> lk(h->mu);
        const int r = (int)(nrhs ? nrhs[i] : 1);
        hipStream_t st = ls[(size_t)lane[(size_t)i]];
        e = matvec_all(*h, x[i], y[i], h->w1, h->w2, r, 1, 0, 0, 0, 0, gla::Epi(), st);
    }
    for (int k = 0; k < lanes && e == hipSuccess; ++k) e = hipEventRecord(ev[(size_t)k], ls[(size_t)k]);
    for (int k = 0; k < lanes && e == hipSuccess; ++k) e = hipStreamWaitEvent(cap, ev[(size_t)k], 0);   // join
    if (capturing) {
        hipGraph_t gr = nullptr;
        const hipError_t ee = hipStreamEndCapture(cap, &gr);
        if (e == hipSuccess) e = ee;
        b->graph = gr;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    for (auto& v : ev)
        if (v) (void)hipEventDestroy(v);
    for (auto& st : ls)
        if (st) (void)hipStreamDestroy(st);
    if (cap) (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return fail(GLA_ERR_HIP, std::string("batch recording: ") + hipGetErrorString(e));
    for (int i = 0; i < n; ++i) {
        std::lock_guard<std::mutex> lk(ops[i]->mu);
        ++ops[i]->batches;
        b->ops.push_back(ops[i]);
    }
    *out = b.release();
    return GLA_OK;
}

gla_status gla_batch_launch(gla_batch_t b, void* stream)
{
    if (!b) return fail(GLA_ERR_ARG, "null batch");
    DevGuard g(b->device);
    const hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    return e == hipSuccess ? GLA_OK : fail(GLA_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
}

gla_status gla_batch_destroy(gla_batch_t b)
{
    delete b;
    return GLA_OK;
}

gla_status gla_opr_reserve(gla_opr_t h, int64_t nrhs)
{
    if (!h) return fail(GLA_ERR_ARG, "null operator");
    if (nrhs < 1 || nrhs > 65535) return fail(GLA_ERR_ARG, "nrhs must be in [1, 65535]");
    std::lock_guard<std::mutex> lk(h->mu);
    DevGuard g(h->opts.device);
    return ensure_ws(h, nrhs);
}

int64_t gla_opr_workspace_bytes(gla_opr_t h, int64_t nrhs)
{
    if (!h) return -1;
    return (h->geom.w1s + h->geom.w2s) * nrhs * (int64_t)elem_bytes(h->dims.dtype);
}

int64_t gla_krylov_dots_workspace(int64_t k, int64_t nv, int64_t n)
{
    if (k < 0 || nv < 0 || n < 0) return -1;
    return gla::krylov_part_doubles(k, nv, n) * 8;
}

static gla_status krylov_check(int32_t dtype, int64_t k, int64_t nv, int64_t n)
{
    if (dtype != GLA_C64 && dtype != GLA_C128) return fail(GLA_ERR_TYPE, "dtype must be GLA_C64 or GLA_C128");
    if (k < 0 || nv < 0 || n < 0) return fail(GLA_ERR_ARG, "negative size");
    if (k > 65535) return fail(GLA_ERR_ARG, "at most 65535 right-hand sides per call");
    if (nv > (1 << 20)) return fail(GLA_ERR_ARG, "too many basis vectors");
    return GLA_OK;
}

gla_status gla_krylov_dots(int32_t dtype, int64_t k, int64_t nv, int64_t n, const void* V, int64_t v_rhs_stride,
                           int64_t v_vec_stride, const void* w, int64_t w_rhs_stride, void* part,
                           int64_t part_bytes, void* h, double* w2, const int64_t* ncol, void* stream)
{
    if (gla_status s = krylov_check(dtype, k, nv, n)) return s;
    if (k == 0 || (nv == 0 && !w2)) return GLA_OK;
    if (n == 0) return fail(GLA_ERR_ARG, "empty vectors");
    if (!w || !part || (nv > 0 && (!V || !h))) return fail(GLA_ERR_ARG, "null argument");
    if (part_bytes < gla_krylov_dots_workspace(k, nv, n))
        return fail(GLA_ERR_SIZE, "partial-sum workspace smaller than gla_krylov_dots_workspace()");
    const hipStream_t st = (hipStream_t)stream;
    const hipError_t e = dtype == GLA_C64
        ? gla::run_krylov_dots<float>(k, (int)nv, n, V, v_rhs_stride, v_vec_stride, w, w_rhs_stride, part, h, w2,
                                      ncol, st)
        : gla::run_krylov_dots<double>(k, (int)nv, n, V, v_rhs_stride, v_vec_stride, w, w_rhs_stride, part, h, w2,
                                       ncol, st);
    return e == hipSuccess ? GLA_OK : fail(GLA_ERR_HIP, hipGetErrorString(e));
}

gla_status gla_krylov_axpy(int32_t dtype, int64_t k, int64_t nv, int64_t n, const void* V, int64_t v_rhs_stride,
                           int64_t v_vec_stride, const void* h, int64_t h_rhs_stride, void* w,
                           int64_t w_rhs_stride, double s, void* part, int64_t part_bytes, double* w2,
                           const int64_t* ncol, void* stream)
{
    if (gla_status st = krylov_check(dtype, k, nv, n)) return st;
    if (k == 0 || n == 0) return GLA_OK;
    if (!w || (nv > 0 && (!V || !h))) return fail(GLA_ERR_ARG, "null argument");
    if (w2 && (!part || part_bytes < gla_krylov_dots_workspace(k, 0, n)))
        return fail(GLA_ERR_SIZE, "a norm needs the partial-sum workspace (gla_krylov_dots_workspace(k, 0, n))");
    const hipStream_t hs = (hipStream_t)stream;
    const hipError_t e = dtype == GLA_C64
        ? gla::run_krylov_axpy<float>(k, (int)nv, n, V, v_rhs_stride, v_vec_stride, h, h_rhs_stride, w,
                                      w_rhs_stride, s, part, w2, ncol, hs)
        : gla::run_krylov_axpy<double>(k, (int)nv, n, V, v_rhs_stride, v_vec_stride, h, h_rhs_stride, w,
                                       w_rhs_stride, s, part, w2, ncol, hs);
    return e == hipSuccess ? GLA_OK : fail(GLA_ERR_HIP, hipGetErrorString(e));
}

gla_status gla_krylov_rows(int32_t dtype, int64_t k, int64_t n, void* V, int64_t v_rhs_stride, int64_t v_vec_stride,
                           const int64_t* idx, void* b, int64_t b_rhs_stride, const double* scale, void* stream)
{
    if (gla_status st = krylov_check(dtype, k, 0, n)) return st;
    if (k == 0 || n == 0) return GLA_OK;
    if (!V || !idx || !b) return fail(GLA_ERR_ARG, "null argument");
    const hipStream_t hs = (hipStream_t)stream;
    const hipError_t e = dtype == GLA_C64
        ? gla::run_krylov_rows<float>(k, n, V, v_rhs_stride, v_vec_stride, idx, b, b_rhs_stride, scale, hs)
        : gla::run_krylov_rows<double>(k, n, V, v_rhs_stride, v_vec_stride, idx, b, b_rhs_stride, scale, hs);
    return e == hipSuccess ? GLA_OK : fail(GLA_ERR_HIP, hipGetErrorString(e));
}

gla_status gla_krylov_givens(int64_t k, int64_t m, const int64_t* js, const double* hcol, int64_t ldh,
                             const double* hn, double* cs, double* sn, double* H, double* g)
{
    if (k < 0 || m < 1) return fail(GLA_ERR_ARG, "k >= 0 and m >= 1 required");
    if (k == 0) return GLA_OK;
    if (!js || !hcol || !hn || !cs || !sn || !H || !g) return fail(GLA_ERR_ARG, "null argument");
    for (int64_t s = 0; s < k; ++s)
        if (js[s] < 0 || js[s] >= m || js[s] >= ldh)
            return fail(GLA_ERR_ARG, "step index outside [0, min(m, ldh))");
    struct Z {
        double r, i;
    };
    auto mul = [](Z a, Z b) { return Z{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; };
    auto add = [](Z a, Z b) { return Z{a.r + b.r, a.i + b.i}; };
    auto scl = [](double c, Z a) { return Z{c * a.r, c * a.i}; };
    auto cnj = [](Z a) { return Z{a.r, -a.i}; };
    std::vector<Z> col((size_t)m + 1);
    for (int64_t s = 0; s < k; ++s) {
        const int64_t j = js[s];
        const Z* hc = (const Z*)hcol + s * ldh;
        Z* snz = (Z*)sn + s * m;
        double* csr = cs + s * m;
        for (int64_t i = 0; i <= j; ++i) col[(size_t)i] = hc[i];
        col[(size_t)j + 1] = Z{hn[s], 0.0};
        // earlier rotations
        for (int64_t i = 0; i < j; ++i) {
            const double c = csr[i];
            const Z z = snz[i], a = col[(size_t)i], b = col[(size_t)i + 1];
            col[(size_t)i] = add(scl(c, a), mul(z, b));
            col[(size_t)i + 1] = add(scl(-1.0, mul(cnj(z), a)), scl(c, b));
        }
        // the rotation that zeroes the subdiagonal entry
        const Z a = col[(size_t)j], b = col[(size_t)j + 1];
        const double aa = std::hypot(a.r, a.i), den = std::hypot(aa, std::hypot(b.r, b.i));
        double c;
        Z z;
        if (den == 0.0) {
            c = 1.0;
            z = Z{0.0, 0.0};
        } else if (aa == 0.0) {
            c = 0.0;
            z = Z{1.0, 0.0};
        } else {
            c = aa / den;
            z = scl(1.0 / den, mul(scl(1.0 / aa, a), cnj(b)));
        }
        csr[j] = c;
        snz[j] = z;
        col[(size_t)j] = add(scl(c, a), mul(z, b));
        col[(size_t)j + 1] = Z{0.0, 0.0};
        Z* Hs = (Z*)H + s * (m + 1) * m;
        for (int64_t i = 0; i <= j + 1; ++i) Hs[i * m + j] = col[(size_t)i];
        Z* gs = (Z*)g + s * (m + 1);
        const Z gj = gs[j];
        gs[j + 1] = scl(-1.0, mul(cnj(z), gj));
        gs[j] = scl(c, gj);
    }
    return GLA_OK;
}

gla_status gla_krylov_next(int32_t dtype, int64_t k, int64_t n, void* V, int64_t v_rhs_stride, int64_t v_vec_stride,
                           const int64_t* idx, const void* w, int64_t w_rhs_stride, const double* w2, void* stage,
                           int64_t stage_rhs_stride, void* stream)
{
    if (gla_status st = krylov_check(dtype, k, 0, n)) return st;
    if (k == 0 || n == 0) return GLA_OK;
    if (!V || !idx || !w || !w2) return fail(GLA_ERR_ARG, "null argument");
    const hipStream_t hs = (hipStream_t)stream;
    const hipError_t e = dtype == GLA_C64
        ? gla::run_krylov_next<float>(k, n, V, v_rhs_stride, v_vec_stride, idx, w, w_rhs_stride, w2, stage,
                                      stage_rhs_stride, hs)
        : gla::run_krylov_next<double>(k, n, V, v_rhs_stride, v_vec_stride, idx, w, w_rhs_stride, w2, stage,
                                       stage_rhs_stride, hs);
    return e == hipSuccess ? GLA_OK : fail(GLA_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"

namespace gla {

bool len_supported(int n)
{
#define GLA_LEN_EQ(k) if (n == k) return true;
    GLA_FOR_EACH_LEN(GLA_LEN_EQ)
#undef GLA_LEN_EQ
    return false;
}

int zpass_tile_width(int complex_bytes, int n, int L1, int L2)
{
    return ztune(complex_bytes, n, (int64_t)L1 * L2 < ZSMALL_COLS).px;
}

std::vector<int> zpass_order(const int B[3], const int L[3], const AxisMap am[3], int px)
{
    const int ntile = (L[0] + px - 1) / px;
    const int nreal = ntile * L[1];
    std::vector<int> ord;
    ord.reserve((size_t)nreal + 8);
    const bool mirror = am[0].sym && am[1].sym && B[0] % px == 0;
    if (!mirror) {
        for (int l = 0; l < nreal; ++l) ord.push_back((l % ntile) | ((l / ntile) << 16));
    } else {
        const int nt1 = B[0] / px;   // tiles per axis-1 parity
        std::vector<char> done((size_t)nreal, 0);
        auto emit = [&](int tile, int m2) {
            const int id = m2 * ntile + tile;
            if (!done[(size_t)id]) { done[(size_t)id] = 1; ord.push_back(tile | (m2 << 16)); }
        };
        for (int p2 = 0; p2 < 2; ++p2)
            for (int k2 = 0; k2 < B[1]; ++k2) {
                const int k2m = p2 == 0 ? (B[1] - k2) % B[1] : B[1] - 1 - k2;
                if (k2m < k2) continue;
                for (int p1 = 0; p1 < 2; ++p1)
                    for (int a = 0; a < nt1; ++a) {
                        const int am_ = nt1 - 1 - a;
                        if (am_ < a) continue;
                        const int t0 = p1 * nt1 + a, t1 = p1 * nt1 + am_;
                        const int r0 = p2 * B[1] + k2, r1 = p2 * B[1] + k2m;
                        emit(t0, r0); emit(t1, r0); emit(t0, r1); emit(t1, r1);
                    }
            }
    }
    while (ord.size() % 8) ord.push_back(-1);
    return ord;
}

// the fast lengths, 2^a x {1, 3, 5}, are the embedding targets; the others
// (29, 81, 145) are compiled for Fourier data that comes in that layout
bool len_fast(int k)
{
    if (k < 1) return false;
    while (k % 2 == 0) k /= 2;
    return (k == 1 || k == 3 || k == 5) && len_supported(k == 1 ? 1 : k);
}

int len_next_supported(int n)
{
    int best = 0;
#define GLA_LEN_GE(k) if (k >= n && len_fast(k) && (best == 0 || k < best)) best = k;
    GLA_FOR_EACH_LEN(GLA_LEN_GE)
#undef GLA_LEN_GE
    return best;
}

}  // namespace gla
