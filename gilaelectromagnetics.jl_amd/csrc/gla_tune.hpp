// Launch shapes per transform length, element type and grid class: the
// tables are the output of gilaelectromagnetics.jl_amd/tune.py (timing of every
// candidate on the target GPU, written between the BEGIN / END markers);
// lengths not listed use the defaults.
//   ztune: pass 3 -- target elements per thread (E, the radix plan follows)
//          and columns per workgroup (PX, also the tile width of the
//          host-built tile order), for grids with many columns and for small
//          cross-sections (L1 L2 < ZSMALL_COLS columns: fewer, narrower
//          workgroups fill the GPU better there)
//   rtune: passes 1, 2, 4, 5 -- elements per thread of the axis-1 rows (ex)
//          and of the axis-2 columns (ey), and the axis-2 bundle width (pxy)
//   xtune: passes 1 and 5 -- the two parities of a row through the stages
//          together (2) or one after the other through half the LDS (1)
//   ytune: passes 2 and 4 -- the two parities of a column bundle through the
//          stages together (2) or one after the other through half the LDS (1)
//   ltune: pass 3's input column -- 0 kept in registers through the even
//          parity, 1 read again (an L2 hit) for the odd parity, 2 both
//          parities split first
//   gtune: pass 3's G^ loads -- 3: root and phase tables in LDS and the G^
//          loads pipelined one point ahead; 7: also the first point's loads
//          issued before the forward transform (more registers)
//   ptune: the radix plan of every pencil transform of that length -- 0 takes
//          the largest radices (fewest LDS exchanges), 1 only radices that
//          divide the elements per thread (every thread busy in every stage,
//          more exchanges)
// The ComplexF32 length-128 pass-3 entry (4, 16) was set from a same-box A/B
// on the 128^3 self matvec (pass 3 161.2 -> 156.9 us over four alternations)
// where the tuner's grid had it 2.8% ahead, inside its 5% margin.
// Tuning builds force one value for all lengths:
//   -DGLA_ZT_E_F=.. -DGLA_ZT_PX_F=.. / -DGLA_ZT_E_D=.. -DGLA_ZT_PX_D=..
//   -DGLA_RT_EX_F=.. -DGLA_RT_EY_F=.. -DGLA_RT_PXY_F=.. (and _D)
//   -DGLA_XK_F=.. / -DGLA_XK_D=..
//   -DGLA_ZGG_F=.. / -DGLA_ZGG_D=..
//   -DGLA_YK_F=.. / -DGLA_YK_D=.. (or -DGLA_YK=.. for both)
//   -DGLA_ZRELOAD_F=.. / -DGLA_ZRELOAD_D=..
//   -DGLA_PT_F=.. / -DGLA_PT_D=..
// (_F: ComplexF32, _D: ComplexF64.)
#pragma once

namespace gla {

constexpr int ZSMALL_COLS = 2 * 128 * 128;

struct ZShape {
    int e, px;
};

struct RShape {
    int ex, ey, pxy;
};

constexpr ZShape ztune_large(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
        switch (n) {
            // BEGIN TUNED C64
            case 16: return {4, 16};
            case 24: return {2, 8};
            case 32: return {4, 8};
            case 40: return {4, 16};
            case 48: return {4, 16};
            case 64: return {4, 8};
            case 80: return {4, 8};
            case 96: return {4, 8};
            case 128: return {4, 16};
            case 160: return {4, 4};
            case 192: return {8, 8};
            case 256: return {4, 8};
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
    }
    switch (n) {
        // BEGIN TUNED C128
            case 16: return {4, 16};
            case 24: return {2, 8};
            case 40: return {2, 8};
            case 96: return {2, 4};
            case 160: return {4, 4};
            case 192: return {4, 4};
            case 256: return {4, 4};
            case 320: return {4, 4};
            case 384: return {4, 4};
            case 512: return {8, 4};
        // END TUNED C128
        default: return {4, 8};
    }
}

// small cross-sections; unlisted lengths take the large-grid shape
constexpr ZShape ztune_small(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
        switch (n) {
            // BEGIN TUNED C64 SMALL
            case 128: return {4, 8};
            case 192: return {4, 4};
            case 384: return {4, 4};
            // END TUNED C64 SMALL
            default: return ztune_large(complex_bytes, n);
        }
    }
    switch (n) {
        // BEGIN TUNED C128 SMALL
            case 48: return {2, 4};
            case 64: return {4, 4};
        // END TUNED C128 SMALL
        default: return ztune_large(complex_bytes, n);
    }
}

constexpr ZShape ztune(int complex_bytes, int n, bool small)
{
#if defined(GLA_ZT_E_F) && defined(GLA_ZT_PX_F)
    if (complex_bytes == 8) return {GLA_ZT_E_F, GLA_ZT_PX_F};
#endif
#if defined(GLA_ZT_E_D) && defined(GLA_ZT_PX_D)
    if (complex_bytes == 16) return {GLA_ZT_E_D, GLA_ZT_PX_D};
#endif
    return small ? ztune_small(complex_bytes, n) : ztune_large(complex_bytes, n);
}

constexpr RShape rtune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_RT_EX_F) && defined(GLA_RT_EY_F) && defined(GLA_RT_PXY_F)
        (void)n;
        return {GLA_RT_EX_F, GLA_RT_EY_F, GLA_RT_PXY_F};
#else
        switch (n) {
            // BEGIN TUNED ROWS C64
            case 48: return RShape{8, 4, 16};
            case 160: return RShape{8, 4, 16};
            // END TUNED ROWS C64
            default: return n <= 160 ? RShape{4, 4, 16} : RShape{8, 8, 16};
        }
#endif
    }
#if defined(GLA_RT_EX_D) && defined(GLA_RT_EY_D) && defined(GLA_RT_PXY_D)
    (void)n;
    return {GLA_RT_EX_D, GLA_RT_EY_D, GLA_RT_PXY_D};
#else
    switch (n) {
        // BEGIN TUNED ROWS C128
            case 48: return RShape{8, 4, 8};
            case 256: return RShape{8, 4, 16};
        // END TUNED ROWS C128
        default: return n <= 128 ? RShape{4, 4, 8} : RShape{8, 8, 8};
    }
#endif
}

constexpr int xtune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_XK_F)
        (void)n;
        return GLA_XK_F;
#else
        switch (n) {
            // BEGIN TUNED XK C64
            case 24: return 2;
            case 80: return 2;
            case 96: return 2;
            case 192: return 2;
            case 320: return 2;
            case 384: return 2;
            // END TUNED XK C64
            default: return 1;
        }
#endif
    }
#if defined(GLA_XK_D)
    (void)n;
    return GLA_XK_D;
#else
    switch (n) {
        // BEGIN TUNED XK C128
            case 80: return 2;
            case 96: return 2;
            case 192: return 2;
            case 320: return 2;
            case 384: return 2;
        // END TUNED XK C128
        default: return 1;
    }
#endif
}

constexpr int gtune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_ZGG_F)
        (void)n;
        return GLA_ZGG_F;
#else
        switch (n) {
            // BEGIN TUNED GG C64
            // END TUNED GG C64
            default: return 7;
        }
#endif
    }
#if defined(GLA_ZGG_D)
    (void)n;
    return GLA_ZGG_D;
#else
    switch (n) {
        // BEGIN TUNED GG C128
            case 320: return 7;
        // END TUNED GG C128
        default: return 3;
    }
#endif
}

constexpr int ytune(int complex_bytes, int n)
{
#if defined(GLA_YK)
    (void)complex_bytes;
    (void)n;
    return GLA_YK;
#else
    if (complex_bytes == 8) {
#if defined(GLA_YK_F)
        (void)n;
        return GLA_YK_F;
#else
        switch (n) {
            // BEGIN TUNED YK C64
            // END TUNED YK C64
            default: return 2;
        }
#endif
    }
#if defined(GLA_YK_D)
    (void)n;
    return GLA_YK_D;
#else
    switch (n) {
        // BEGIN TUNED YK C128
            case 512: return 1;
        // END TUNED YK C128
        default: return 2;
    }
#endif
#endif
}

constexpr int ltune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_ZRELOAD_F)
        (void)n;
        return GLA_ZRELOAD_F;
#else
        switch (n) {
            // BEGIN TUNED RELOAD C64
            // END TUNED RELOAD C64
            default: return 0;
        }
#endif
    }
#if defined(GLA_ZRELOAD_D)
    (void)n;
    return GLA_ZRELOAD_D;
#else
    switch (n) {
        // BEGIN TUNED RELOAD C128
        // END TUNED RELOAD C128
        default: return 2;
    }
#endif
}

constexpr int ptune(int complex_bytes, int n)
{
    if (complex_bytes == 8) {
#if defined(GLA_PT_F)
        (void)n;
        return GLA_PT_F;
#else
        switch (n) {
            // BEGIN TUNED PLAN C64
            // END TUNED PLAN C64
            default: return 0;
        }
#endif
    }
#if defined(GLA_PT_D)
    (void)n;
    return GLA_PT_D;
#else
    switch (n) {
        // BEGIN TUNED PLAN C128
        // END TUNED PLAN C128
        default: return 0;
    }
#endif
}

}  // namespace gla
