/*
 * gmcore_ens.h -- batched ensembles of shallow-water contexts on one MI355X.
 *
 * An ensemble batches already prepared gmc_ctx members (gmcore_hip.h), of any
 * resolutions, so that one step of every member is one HIP graph of three
 * launches per stage: the stage kernels of gmc_step, with the tiles and
 * filter work items of all members numbered consecutively.  Member k steps
 * with its own dt[k]; each member gets exactly the bits gmc_step would give
 * it.  The reference runs ensemble members as separate jobs; the per-member
 * semantics are those of gmc_step (time_integrator + damp_run + swap,
 * time_schemes_mod.F90:244-288, gmcore_mod.F90:138-144).
 *
 * Conventions are those of gmcore_hip.h.
 */
#ifndef GMCORE_ENS_H
#define GMCORE_ENS_H

#include <stdint.h>
#include "gmcore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmc_ens gmc_ens;

/* Batch n contexts.  Every member must have run gmc_prepare, use the same
 * time scheme, be at the same old level and have divergence damping off.
 * The members keep ownership of their fields; while the ensemble exists they
 * are stepped only through it (gmc_download and gmc_diagnose stay valid
 * after gmc_ens_step returns). */
int gmc_ens_create(gmc_ctx *const *members, int32_t n, gmc_ens **out);
int gmc_ens_destroy(gmc_ens *ens);
const char *gmc_ens_last_error(const gmc_ens *ens);

/* nsteps steps of every member, member k with time step dt[k]; replays
 * captured graphs (per old/new parity: of 8 steps while 8 or more remain,
 * then of one step) and synchronises at the end. */
int gmc_ens_step(gmc_ens *ens, const double *dt, int32_t nsteps);

/* Launch shapes: the rows of the tendency kernel's tile and the size of its
 * persistent grid.  gmc_ens_autotune times trial_steps steps of every shape
 * on a snapshot of the members (restored afterwards, so the state does not
 * advance), selects the fastest and returns its index in *choice; times_ms[k]
 * (up to cap entries) receives the mean step time of shape k and *ncand the
 * number of shapes.  Every shape gives the same bits. */
int gmc_ens_autotune(gmc_ens *ens, const double *dt, int32_t trial_steps, int32_t *choice,
                     double *times_ms, int32_t cap, int32_t *ncand);
int gmc_ens_shape(gmc_ens *ens, int32_t shape);
const char *gmc_ens_shape_name(int32_t shape);

/* The HIP stream of the ensemble's launches. */
int gmc_ens_stream(gmc_ens *ens, void **stream);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_ENS_H */
