// Green-tensor setup (operator construction from geometry).
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "gila_amd.h"

namespace gla {

// Fourier Green data of the operator (trg <- src) in the reference's 8-branch
// layout (src/glaMemSup.jl:138-156), computed in FP64.
gla_status setup_fur(const gla_opts& opts, const gla_vol& trg, const gla_vol& src, const gla_dims& d,
                     std::vector<std::complex<double>>& fur, std::string& err);

}  // namespace gla
