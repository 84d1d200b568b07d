/*
 * nbody_mlp_oracle.c -- CPU oracle for the learned acceleration extrapolator.
 *
 * TEST INFRASTRUCTURE ONLY (see nbody_oracle.c).  No reference counterpart:
 * the reference has no learned component.  Pinned by the tests against
 * closed forms (polynomial extrapolation is exact on cubics) and against
 * central finite differences of the FP64 loss (gradients).
 */
#include <math.h>
#include <stddef.h>

#define ORC_API __attribute__((visibility("default")))

/* Cubic extrapolation from the last four equally spaced samples:
 * a_{n+1} = 4 a_n - 6 a_{n-1} + 4 a_{n-2} - a_{n-3} (exact for cubics). */
ORC_API void orc_poly_extrap(size_t n3, const double *a0, const double *a1, const double *a2,
                             const double *a3, double *out)
{
    for (size_t i = 0; i < n3; i++) out[i] = 4.0 * a0[i] - 6.0 * a1[i] + 4.0 * a2[i] - a3[i];
}
