// Host build of the exact Stokes step (bh_stokes_exact.h) for the CPU tests; test infrastructure only.
#include "bh_stokes_exact.h"

extern "C" {

// Halvings exact_operator takes for (c, ds); -1 for a non-finite alpha ds.
int bh_test_exact_halvings(const double *c, double ds) {
    double A[4][4], b[4];
    int p = 0;
    bh::exact_operator(c, ds, A, b, p);
    return p;
}

void bh_test_stokes_exact(const double *c, double ds, double *s) { bh::stokes_exact(s, c, ds); }

void bh_test_stokes_exact_f32(const float *c, float ds, float *s) { bh::stokes_exact(s, c, ds); }

}  // extern "C"
