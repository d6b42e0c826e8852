/*
 * fmd_ext.h -- extensions of the public API (libfmd.so).
 *
 * None of these has a counterpart in the reference's src/fmd.h.  They exist
 * for callers that need binary state exchange or direct access to the
 * device system behind an fmd_t: benchmarks, tests and batch drivers.
 */
#ifndef FMD_EXT_H
#define FMD_EXT_H

#include "fmd.h"
#include "fmdx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Append n atoms, in the given order, to the global grid: the binary form of
   fmd_io_loadState (same binning by floor(x/cellh) and the same within-cell
   order as loading a state file written in this order).  x, v: [3n]. */
void fmd_ext_loadAtoms(fmd_t *md, unsigned n, const double *x, const double *v, const int *group,
                       const unsigned *kind);

/* Upload the atoms and potentials to the GPU now (otherwise done by the
   first fmd_dync_* call).  Returns false on failure. */
bool fmd_ext_distribute(fmd_t *md);

/* The device system of an instance after distribution (NULL before). */
fmdx_system_t *fmd_ext_getDevice(fmd_t *md);

/* Force-kernel arithmetic: FMDX_PREC_FP64 (default) or FMDX_PREC_MIXED. */
void fmd_ext_setPrecision(fmd_t *md, int precision);

/* Atoms in cell order (the order of fmd_io_saveState); any pointer may be
   NULL; returns the atom count.  Call with all pointers NULL to size buffers. */
unsigned fmd_ext_getAtoms(fmd_t *md, double *x, double *v, double *f, int *group, unsigned *atomid,
                          unsigned *kind);

/* Potential energy of the last force evaluation (eV). */
double fmd_ext_getPotentialEnergy(fmd_t *md);

/* TTM turies beyond 1 x 1 x n (fmd_turi_add with dimx or dimy > 1, type 1):
   the reference accepts them but has no updater for them (ttm.c:451-452).
   This build solves them on the device, periodic in x and y.
   Conductivity kappa = K0 * Te / Tl (K0 in W/m/K) instead of a constant: */
void fmd_ext_ttm_setHeatConductivity_proportional(fmd_t *md, fmd_handle_t turi, fmd_real_t K0);
/* Solve a 1 x 1 x n type-1 turi with the grid solver (used to show that the
   grid solver reduces to the reference's 1-D scheme). */
void fmd_ext_ttm_useGridSolver(fmd_t *md, fmd_handle_t turi, bool on);

#ifdef __cplusplus
}
#endif

#endif /* FMD_EXT_H */
