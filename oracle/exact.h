/*
 * ORACLE — test infrastructure only. Nothing in the product links this file.
 *
 * Exact sign of 3x3 / 4x4 determinants of double entries, by expansion
 * arithmetic (Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
 * Robust Geometric Predicates", DCG 1997: Two-Sum, Two-Product,
 * Grow-Expansion), behind a static forward-error filter.  Used as the
 * conflict predicate of the oracle's convex cell, the way geogram's
 * VBW::ConvexCell uses PCK::det_4d (numerics/predicates.cpp:2073-2110,
 * numerics/predicates/det4d.h) when exact predicates are on
 * (voronoi/convex_cell.cpp:137, 870-911).
 */
#ifndef VGC_ORACLE_EXACT_H
#define VGC_ORACLE_EXACT_H

#ifdef __cplusplus
extern "C" {
#endif

/* sign (-1, 0, +1) of det[r0; r1; r2; r3] (rows of 4 doubles), exact. */
int orc_det4_sign(const double* r0, const double* r1, const double* r2,
                  const double* r3);

/* sign of det[r0; r1; r2] (rows of 3 doubles), exact. */
int orc_det3_sign(const double* r0, const double* r1, const double* r2);

/* same as orc_det4_sign but always through expansions (no filter). */
int orc_det4_sign_exact(const double* r0, const double* r1, const double* r2,
                        const double* r3);

/* number of calls that fell through the filter since start (statistics). */
long long orc_exact_fallbacks(void);

#ifdef __cplusplus
}
#endif
#endif
