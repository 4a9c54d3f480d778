// Launch layer between the C-ABI (gla_capi.cpp) and the pass kernels.
#pragma once
#include "gla_tune.hpp"
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>
#include "gla_kernels.hpp"

// Branch lengths B (per-axis FFT lengths) compiled into the device library:
// 2^a 3^b 5^c lengths up to 512, and 29, 81 and 145 (the branch lengths of a
// 256x34x34 source in a 256x256x128 target, so that Fourier data in the
// reference's layout for that pair is accepted as it is).  Operators built
// from geometry reach other lengths by embedding into the next listed length
// (the circulant only needs L >= S + T - 1).
#define GLA_FOR_EACH_LEN(X)                                                                          \
    X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(29) X(32) X(40) X(48) X(64) \
        X(80) X(81) X(96) X(128) X(145) X(160) X(192) X(256) X(320) X(384) X(512)

namespace gla {

struct Geom {
    int S[3], T[3], B[3], L[3];   // partition cells (the whole volume without partitions)
    int Y1, Z;             // padded workspace extents: max(S2,T2), max(S3,T3)
    int64_t xs, ys;        // per-RHS strides of X and Y (elements, whole volumes)
    int64_t w1s, w2s;      // per-RHS strides of W1 and W2 (elements)
    Part xp, yp;           // source / target partition being applied
};

// the whole volume of n[0] x n[1] x n[2] cells as its only partition
inline Part whole_part(const int* n)
{
    Part p{};
    for (int a = 0; a < 3; ++a) { p.f[a] = n[a]; p.d[a] = 1; p.o[a] = 0; }
    return p;
}

struct DevTab {
    const void* phz[3];    // phi_d[n] = e^{-i pi n / B_d}, n < B_d
    const void* tw[3];     // e^{-2 pi i n / B_d}, n < B_d
    AxisMap am[3];
    const void* gf;        // G^ in device layout
    const int* zorder;     // pass-3 tile order (tile | m2 << 16, -1 padding)
    int zper_xcd;          // pass-3 workgroups per XCD (grid = 8 * zper_xcd)
};

// columns per pass-3 workgroup for the element type (bytes per complex), the
// pass-3 transform length and the grid's L1 x L2 columns (the tables of
// gla_tune.hpp)
int zpass_tile_width(int complex_bytes, int n, int L1, int L2);

// Pass-3 tile order: ntile = ceil(L1 / PX) column tiles times L2 rows,
// padded to a multiple of 8 with -1.  With symmetric storage on axes 1 and 2
// the tiles whose stored Green entries coincide (mirror images) are listed
// next to each other.
std::vector<int> zpass_order(const int B[3], const int L[3], const AxisMap am[3], int px);


// last-pass epilogue operands for y = x - chi (G x) (axpby mode 2)
struct Epi {
    const void* x = nullptr;   // source field (self operator), per-RHS stride xs
    int64_t xs = 0;
    const void* chi = nullptr; // one value per target cell, per-RHS stride chis (0: shared)
    int64_t chis = 0;
};

bool len_supported(int n);
bool len_fast(int n);             // 2^a x {1, 3, 5}: the lengths operators from geometry embed into
int len_next_supported(int n);   // smallest fast length >= n, or 0

constexpr int NPASS = 5;

// one pass of the pipeline: 0 x-forward, 1 y-forward, 2 z fused, 3 y-inverse,
// 4 x-inverse
template <typename R>
hipError_t run_pass(int pass, const Geom& g, const DevTab& tb, const void* X, void* Y, void* W1, void* W2,
                    int nrhs, double ar, double ai, double br, double bi, int axpby, const Epi& epi,
                    hipStream_t st);

// forward x, y, z passes only (operator setup): X (S1,S2,S3,3) -> W2 with
// every axis split and transformed
template <typename R>
hipError_t run_fwd3(const Geom& g, const DevTab& tb, const void* X, void* W1, void* W2, int nrhs, hipStream_t st);

template <typename R>
hipError_t run_matvec(const Geom& g, const DevTab& tb, const void* X, void* Y, void* W1, void* W2,
                      int nrhs, double ar, double ai, double br, double bi, int axpby, const Epi& epi,
                      hipStream_t st);

}  // namespace gla
