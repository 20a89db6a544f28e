/* TEST INFRASTRUCTURE ONLY (the oracle).  Never linked into the product.
 *
 * Sequential CPU restatement of the reference's time-marching loop
 * (multall-open-18.3.f: LOOP :3664-5463, TSTEP :5470-5932, SMOOTH_VAR
 * :15352-15594, SMOOTH_RESID :14833-14917, STEPUP :7916-8002, NEW_MIXPLAN
 * :14922-15251, LOSS :8058-9066, SPAL_LOSS :13586-14827).  Arithmetic is FP32 in the reference's
 * statement order so the result can be compared bit-for-bit with the
 * compiled reference on the same host.
 *
 * Storage is the reference's I-fastest column-major order, densely packed to
 * the logical extents: a node/cell array is [KM][JM][IM], a (J,K) array is
 * [KM][JM], an (I,K) array is [KM][IM].  Integer maps keep the reference's
 * 1-based values.  The caller owns every buffer and binds it by name.
 */
#ifndef MH_ORACLE_H
#define MH_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t *orc_create(void);
void orc_destroy(orc_t *o);
/* Bind a caller-owned array by its reference name (e.g. "RO", "JB1").
 * Returns 0 on success, -1 for an unknown name. */
int orc_bind(orc_t *o, const char *name, void *ptr);
/* Set a scalar by its reference name.  Returns -1 for an unknown name. */
int orc_scalar(orc_t *o, const char *name, double value);
double orc_get_scalar(orc_t *o, const char *name);
/* Check bindings and perform the once-per-run set-up at the top of LOOP
 * (multall-open-18.3.f:3674-3714).  Returns 0 or a negative error code. */
int orc_start(orc_t *o);
/* Advance n time steps (the body of LOOP from label 5000).  Returns the
 * number of steps run; stops early if the reference would (IFEND). */
int orc_step(orc_t *o, int n);
/* Individual phases, exposed for kernel-level comparisons. */
void orc_primitives(orc_t *o);
void orc_tstep(orc_t *o, float *d, float *diff, int ncall);
void orc_smooth_var(orc_t *o, float *d);
void orc_stepup(orc_t *o, float relax);
void orc_new_mixplan(orc_t *o);
void orc_loss(orc_t *o);
void orc_spal_loss(orc_t *o);
void orc_new_loss(orc_t *o);

#ifdef __cplusplus
}
#endif
#endif
