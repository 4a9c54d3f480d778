// Pass-3 shape per transform length and element type: the target elements per
// thread (E, the radix plan follows from it) and the columns per workgroup
// (PX, also the tile width of the host-built tile order).  The table is the
// result of gilaelectromagnetics.jl_amd/tune_zpass.py (timing of every
// candidate shape at every length on the target GPU); lengths not listed use
// the defaults.  Tuning builds force one shape for all lengths with
// -DGLA_ZT_E_F=.. -DGLA_ZT_PX_F=.. (ComplexF32) / -DGLA_ZT_E_D=.. -DGLA_ZT_PX_D=..
// The ComplexF32 entry for 128 was re-measured after the pass-3 addressing
// change (same-GPU timing of the 128^3 operator: {4, 8} 161.6 us against
// 178.2 us for the default {8, 16}).
#pragma once

namespace gla {

struct ZShape {
    int e, px;
};

constexpr ZShape ztune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_ZT_E_F) && defined(GLA_ZT_PX_F)
        (void)n;
        return {GLA_ZT_E_F, GLA_ZT_PX_F};
#else
        switch (n) {
            // BEGIN TUNED C64
            case 16: return {4, 16};
            case 20: return {4, 8};
            case 24: return {4, 16};
            case 32: return {4, 8};
            case 40: return {4, 16};
            case 48: return {4, 16};
            case 80: return {4, 8};
            case 96: return {4, 8};
            case 128: return {4, 8};
            case 160: return {4, 16};
            case 192: return {8, 8};
            case 256: return {8, 8};
            case 320: return {4, 8};
            case 384: return {4, 8};
            case 512: return {8, 8};
            // END TUNED C64
            // lengths with a prime factor 29 (29 threads per column): 8
            // columns per workgroup keep its 232 threads within one wave per
            // SIMD, so the radix-29 butterflies get the full register file
            case 29: return {1, 8};
            case 145: return {5, 8};
            default: return {8, 16};
        }
#endif
    }
#if defined(GLA_ZT_E_D) && defined(GLA_ZT_PX_D)
    (void)n;
    return {GLA_ZT_E_D, GLA_ZT_PX_D};
#else
    switch (n) {
        // BEGIN TUNED C128
            case 20: return {2, 8};
            case 64: return {4, 4};
            case 160: return {4, 4};
            case 192: return {4, 4};
            case 256: return {8, 4};
            case 320: return {4, 4};
            case 384: return {4, 4};
            case 512: return {8, 4};
        // END TUNED C128
        default: return {4, 8};
    }
#endif
}

}  // namespace gla
