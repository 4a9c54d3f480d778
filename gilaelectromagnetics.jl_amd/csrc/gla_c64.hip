// ComplexF32 instantiation of the matvec passes.
#include "gla_launch_impl.hpp"
template hipError_t gla::run_matvec<float>(const gla::Geom&, const gla::DevTab&, const void*, void*, void*, void*,
                                           int, double, double, double, double, int, const gla::Epi&, hipStream_t);
template hipError_t gla::run_pass<float>(int, const gla::Geom&, const gla::DevTab&, const void*, void*, void*, void*,
                                      int, double, double, double, double, int, const gla::Epi&, hipStream_t);
