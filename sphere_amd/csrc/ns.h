// ns.h -- the Navier-Stokes fluid step as the DEM step calls it when a
// fluid is attached to a context (sph_ns_attach, ns.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/sphere_hip.h"

// One DEM iteration of the fluid with the context's grains (original-order
// x and vel rows of 4, the cell sort's sorted x, cell starts and ends and
// original indices), enqueued on stream s: porosities, the interaction
// forces (each grain's sum added to force, rows of 3) and, every ndem
// iterations, the solve (src/device.cu:1213-1898); wall_x >= 0 is the host
// copy of a top wall's position (wmode 1 or 3), < 0 none.  0 or an SPH_ERR_*
// code.
int ns_coupled_step(sph_ns *ns, uint32_t np, const double *x, const double *vel, const double *xs,
                    const uint32_t *cstart, const uint32_t *cend, const uint32_t *gidx,
                    const double *origo, int periodic, double *force, double wall_x, hipStream_t s);
// The fluid's grid, to check that it is the context's grid.
void ns_grid(const sph_ns *ns, uint32_t num[3], double L[3]);
// Marks the fluid attached to a context (1) or free (0).
void ns_set_attached(sph_ns *ns, int on);
// An output event of the attached context: the host copy of p that the
// fluid's boundary guesses read is refreshed, as the reference's transfer
// for output does (src/device.cu:2576-2596).
int ns_output_event(sph_ns *ns, hipStream_t s);
