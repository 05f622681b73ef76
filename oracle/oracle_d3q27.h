// ORACLE — test infrastructure only. Not part of the product; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// D3Q27 lattice tables for the CPU restatement.
// Direction order and inverse pairs follow src/lbm/constants/D3Q27.h:46-110.
// Weights follow src/basics/constants/NumericConstants.h:79-85 (8/27, 2/27, 1/54, 1/216).
#pragma once

namespace orc {

enum Dir : int {
    d000 = 0, dP00, dM00, d0P0, d0M0, d00P, d00M,
    dPP0, dMM0, dPM0, dMP0, dP0P, dM0M, dP0M, dM0P, d0PP, d0MM, d0PM, d0MP,
    dPPP, dMPP, dPMP, dMMP, dPPM, dMPM, dPMM, dMMM
};

// lattice velocity components (src/lbm/constants/D3Q27.h:46-72 names encode them)
static const int CX[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1};
static const int CY[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1,-1};
static const int CZ[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1, 1, 1, 1, 1,-1,-1,-1,-1};

// opposite direction (src/lbm/constants/D3Q27.h:84-110)
static const int INV[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
                            26, 25, 24, 23, 22, 21, 20, 19};

inline int dir_of(int cx, int cy, int cz)
{
    for (int d = 0; d < 27; ++d)
        if (CX[d] == cx && CY[d] == cy && CZ[d] == cz) return d;
    return -1;
}

template <class T> inline T weight(int d)
{
    const int s = CX[d] * CX[d] + CY[d] * CY[d] + CZ[d] * CZ[d];
    if (s == 0) return T(8.) / T(27.);
    if (s == 1) return T(2.) / T(27.);
    if (s == 2) return T(1.) / T(54.);
    return T(1.) / T(216.);
}

} // namespace orc
