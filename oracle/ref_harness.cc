// Thin C-ABI harness over the reference's own generated serial operators.
// TEST INFRASTRUCTURE ONLY: it is compiled by oracle/Makefile directly from
// the reference tree (Cxx11/stencil_seq.hpp, Cxx11/prk_util.h, included
// where they lie) into oracle/_ref/libprkref.so, and the tests use it to pin
// the oracle's restatement bit for bit.  Nothing here is the reference's
// code: this file only adapts raw pointers to the reference's signature
// void starR(const int n, const int t, prk::vector<double>&, prk::vector<double>&)
// (Cxx11/stencil_seq.hpp:16) and reproduces the operator switch of
// Cxx11/stencil.cc:142-159.
#include "prk_util.h"
#include "stencil_seq.hpp"

extern "C" int prkref_stencil(int pattern, int radius, int n, int t, const double *in, double *out)
{
    using fn = void (*)(const int, const int, prk::vector<double> &, prk::vector<double> &);
    fn f = nullptr;
    if (pattern == 0) {
        switch (radius) {
        case 1: f = star1; break;
        case 2: f = star2; break;
        case 3: f = star3; break;
        case 4: f = star4; break;
        case 5: f = star5; break;
        }
    } else {
        switch (radius) {
        case 1: f = grid1; break;
        case 2: f = grid2; break;
        case 3: f = grid3; break;
        case 4: f = grid4; break;
        case 5: f = grid5; break;
        }
    }
    if (!f) return -1;
    const size_t nn = (size_t)n * (size_t)n;
    prk::vector<double> vin(const_cast<double *>(in), nn);
    prk::vector<double> vout(out, nn);
    f(n, t, vin, vout);
    return 0;
}
