/*
 * gdon3d-hip: MI355X-native P2 DG Maxwell hot path behind gdon3d's C API.
 *
 * Precision switch. Replaces the reference's src/gdon3d.h:20-56. The
 * reference defaults to float unless _DOUBLE_PRECISION is defined and its
 * build always defines it (CMakeLists.txt:26); this library is built for
 * double only, so gdn_real is double unconditionally and every struct of the
 * API keeps the reference's double-precision layout. Reduced-precision field
 * storage is a run-time option of the GPU solver
 * (dg_stepper_create_precision),
 * not a change of gdn_real.
 */
#ifndef GDON3D_H
#define GDON3D_H

#ifdef __cplusplus
extern "C" {
#endif

#include <float.h>
#include <limits.h>

#ifndef _DOUBLE_PRECISION
#define _DOUBLE_PRECISION
#endif

#define gdn_real double
#define gdn_value double

#define _VERY_SMALL (1e-12)
#define _SMALL (1e-6)

#define _GDN_XREF_BOUND_THRESHOLD DBL_EPSILON
#define _GDN_XREF_UPBOUND (1.0 + _GDN_XREF_BOUND_THRESHOLD)
#define _GDN_XREF_LOWBOUND (-_GDN_XREF_BOUND_THRESHOLD)

#ifdef __cplusplus
}
#endif
#endif
