/*
 * oracle.h — TEST INFRASTRUCTURE.  CPU restatement of the reference's
 * explicit step (Liangerty/CombustionCodeGPU), used only as the checker by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing
 * in the product links or calls it.
 *
 * The oracle takes the same descriptor structs as the product's C ABI
 * (include/cfx.h) so both sides see identical inputs, and mirrors its entry
 * points with an orc_ prefix.  All arithmetic is FP64 and follows the
 * reference's source text; the functions cite the file:line they restate.
 */
#ifndef CFX_ORACLE_H
#define CFX_ORACLE_H

#include "../include/cfx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int orc_create(const cfx_params *params, const cfx_species *species,
               const cfx_block *blocks, int n_block, orc_ctx **out);
void orc_destroy(orc_ctx *ctx);
const char *orc_last_error(const orc_ctx *ctx);
int orc_set_threads(int n);

int orc_upload_primitive(orc_ctx *ctx, int blk, const double *bv, const double *yk);
int orc_upload_state(orc_ctx *ctx, int blk, const double *cv, const double *bv, const double *yk);
int orc_download(orc_ctx *ctx, int blk, int what, double *dst);
int orc_init_conservative(orc_ctx *ctx);
int orc_store_last_step(orc_ctx *ctx);
int orc_rhs(orc_ctx *ctx, int blk);
int orc_time_step(orc_ctx *ctx, int global, double *dt_out);
int orc_update(orc_ctx *ctx, int stage);
int orc_lusgs(orc_ctx *ctx);
/* Test helper: overwrite dq of block blk (n_var x interior). */
int orc_set_dq(orc_ctx *ctx, int blk, const double *dq);
int orc_apply_bc(orc_ctx *ctx);
int orc_exchange(orc_ctx *ctx);
int orc_eliminate_k_gradient(orc_ctx *ctx);
int orc_update_properties(orc_ctx *ctx);
int orc_residual(orc_ctx *ctx, double res[4]);
int orc_step(orc_ctx *ctx, int n_steps);
int orc_get_dt(orc_ctx *ctx, double *dt);
int orc_set_dt(orc_ctx *ctx, double dt);
long long orc_newton_failures(orc_ctx *ctx);

/* Point-wise building blocks, exposed for known-answer tests. */
/* Thermo.cu:27-47 — h [J/kg] and cp [J/(kg K)] of every species at t. */
int orc_enthalpy_cp(const cfx_species *sp, double t, double *h, double *cp);
/* Transport.cu:43-96 — mu, kappa, rho*D (n_spec) at (t, Y) with cp from
 * compute_cp; returns mixture mw through *mw_out. */
int orc_transport(const cfx_species *sp, double schmidt, double t, const double *Y,
                  double *mu, double *kappa, double *rhoD);
/* Thermo.cu:89-125 — Newton inversion of T from e; returns iterations used. */
int orc_newton_temperature(const cfx_species *sp, int quirks, const double *Y,
                           double e, double t_guess, double *t_out);
/* Limiter.cuh:13-16 (tag 0) and the Van Albada limiter (tag 1). */
double orc_limiter(int tag, double a, double b);
/* Transport.cu:7-10 */
double orc_sutherland(double t);
/* One AUSM+ face flux (InviscidScheme.cu:24-108, Reconstruction.cu:46-89)
 * from four points of primitive data pv[4][5+ns] (rho,u,v,w,p,Y...), the
 * metric row k of the face direction at the two face points
 * (metric[2][3]) and jacobians jac[2].  flux has 5+ns entries. */
int orc_inviscid_face(const cfx_species *sp, const cfx_params *prm,
                      const double *pv, const double *metric, const double *jac,
                      double *flux);

/* The face flux the step differences (inviscid_face_at / viscous_face_at):
 * the face between interior point (i,j,k) and (i,j,k) + e_dir of block blk
 * (i may be -1 along dir), pass 0 = AUSM+ flux, 1 = viscous flux; out has
 * 5+ns entries.  Uses the context's current bv, yk, metric and properties. */
int orc_face_flux(orc_ctx *ctx, int blk, int dir, int i, int j, int k, int pass, double *out);

#ifdef __cplusplus
}
#endif
#endif
