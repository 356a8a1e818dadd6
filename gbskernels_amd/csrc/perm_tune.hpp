// perm_tune.hpp -- built-in Gray-code chunk shape of the permanent kernels:
// log2 of the Gray terms one lane walks, per matrix side n = 0..28, for the
// fp64 family (fp64 / kappa / certified) and the dd family (dd / certified-dd).
// The kernels cap it at the one-workgroup-per-matrix limit; below that limit
// a matrix spreads over 2^(limit - value) workgroups. gbskernels_amd/autotune.py
// measures the candidates on the device and rewrites this table; these entries
// were measured on one MI355X (gfx950) with `python -m gbskernels_amd.autotune
// --write` (fp64 n = 28 and dd n = 14, 24..28 moved off the defaults 9 / 6).
#pragma once

namespace gbsk {

constexpr int kPermLaneLog2[2][29] = {
    // fp64 family
    {9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 10},
    // dd family
    {6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 8, 9, 9, 11},
};

}  // namespace gbsk
