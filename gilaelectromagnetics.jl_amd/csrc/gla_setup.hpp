// Green-tensor setup (operator construction from geometry).
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "gila_amd.h"

namespace gla {

// The 18 singular face-pair values [wS(3), wE(9), wV(6)] of a cell of sizes
// s (wavelengths), divided by the cell volume, at Gauss-Legendre order ord:
// wekS / wekE / wekV (src/glaInt.jl:75-267) as egoFunSng! uses them
// (src/glaGen.jl:117-121).  Runs on the current device.
gla_status singular_values(const double s[3], double zr, double zi, int ord, std::complex<double> w18[18],
                           std::string& err);

// Fourier data in the reference layout at branch lengths dx.brn that have no
// compiled transform, re-expressed at the compiled lengths df.dev_brn: per
// partition pair the full spectrum is expanded from the stored branches,
// transformed back to the real-space circulant (FP64, direct DFT on the GPU),
// re-embedded on the L' = 2 df.dev_brn grid (the circulant only needs
// L' >= S + T - 1) and transformed forward there; out holds the result in the
// reference layout of df.  Runs on the current device.
gla_status refur(const gla_dims& dx, const gla_dims& df, const std::vector<std::complex<double>>& in,
                 std::vector<std::complex<double>>& out, std::string& err);

// Fourier Green data of the operator (trg <- src) in the reference's 8-branch
// layout (src/glaMemSup.jl:138-156), computed in FP64.
gla_status setup_fur(const gla_opts& opts, const gla_vol& trg, const gla_vol& src, const gla_dims& d,
                     std::vector<std::complex<double>>& fur, std::string& err);

}  // namespace gla
