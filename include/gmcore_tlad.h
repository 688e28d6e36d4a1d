/*
 * gmcore_tlad.h -- tangent-linear and adjoint of the shallow-water step on
 * one MI355X, with trajectory checkpointing, for gradients of a cost on the
 * initial state.
 *
 * The reference has no tangent-linear or adjoint model; these entry points
 * differentiate its step: time_integrator with step_forward_backward
 * (time_schemes_mod.F90:102-118, 244-288) of the space operators and
 * update_state (gmcore_mod.F90:343-575, time_schemes_mod.F90:120-242), with
 * divergence damping off.  A gmc_tlad object holds the perturbation (TL) or
 * adjoint (AD) variables of the three state levels of one context
 * (gmcore_hip.h): u, v, gz and the filtered u_f, v_f.  The trajectory is the
 * context's own state, advanced exactly as gmc_step advances it.
 *
 * Non-smooth points of the step (|t| and sign(t) in the upwind pv weights)
 * take the derivative of the branch the trajectory took.  Conventions are
 * those of gmcore_hip.h.
 */
#ifndef GMCORE_TLAD_H
#define GMCORE_TLAD_H

#include <stdint.h>
#include "gmcore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum gmc_tlad_var { GMC_TLAD_U = 0, GMC_TLAD_V, GMC_TLAD_GZ, GMC_TLAD_UF, GMC_TLAD_VF };

typedef struct gmc_tlad gmc_tlad;

/* Attach to a prepared context (gmc_prepare done, divergence damping off).
 * All variables start at zero. */
int gmc_tlad_create(gmc_ctx *ctx, gmc_tlad **out);
int gmc_tlad_destroy(gmc_tlad *t);
const char *gmc_tlad_last_error(const gmc_tlad *t);

/* Perturbation / adjoint variable `var` of state level `level` (1..3) <->
 * host (interior, rows as the matching state field; pole rows of u unused). */
int gmc_tlad_upload(gmc_tlad *t, int32_t level, int32_t var, const double *host, int64_t ld);
int gmc_tlad_download(gmc_tlad *t, int32_t level, int32_t var, double *host, int64_t ld);
int gmc_tlad_zero(gmc_tlad *t);

/* Tangent-linear model: nsteps steps of the context (as gmc_step: the state
 * advances and old/new swap) and of the perturbation (u, v, gz) of its
 * current old level, which then sits on the new current level. */
int gmc_tl_step(gmc_tlad *t, double dt, int32_t nsteps);

/* Adjoint model: on entry the adjoint (u, v, gz) of the state nsteps steps
 * after the context's current old state sits on the current old level; on
 * return that level holds the adjoint of the current state (M^T applied) and
 * the other adjoint levels are zero.  The trajectory is checkpointed every 8
 * steps and each segment's tapes recomputed once; the context's state is
 * left where it was. */
int gmc_ad_step(gmc_tlad *t, double dt, int32_t nsteps);

/* Cost and gradient over a window of nsteps steps from the context's current
 * state x0:
 *   J = 1/2 sum_k (gz_N(cell_k) - y_k)^2 / sigma_o^2 + 1/2 |x0 - xb|^2 / sigma_b^2
 * (gz_N: gz after nsteps steps; cell_k = r * nlon + i, distinct; |.| over u on
 * the rows between the poles, v and gz).  Checkpoints of the trajectory every
 * ckpt_every steps; during the reverse sweep each segment runs forward once
 * from its checkpoint, keeping every stage's tape (24 fields of the grid per
 * stage and step, 26 with coriolis_scheme 2, in two buffers: memory grows
 * with ckpt_every), then its adjoint replays from the tapes.  Writes J and dJ/du0, dJ/dv0, dJ/dgz0 (host, leading dimension ld);
 * the context's state is left at x0 and every adjoint variable at zero. */
int gmc_gradient(gmc_tlad *t, double dt, int32_t nsteps, int32_t ckpt_every, int32_t nobs,
                 const int64_t *obs_cell, const double *obs_value, double sigma_o, const double *u_b,
                 const double *v_b, const double *gz_b, double sigma_b, int64_t ld, double *cost,
                 double *grad_u, double *grad_v, double *grad_gz);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_TLAD_H */
