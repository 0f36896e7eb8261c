/*
 * Oracle: specified (and nested) lateral boundaries.  TEST INFRASTRUCTURE.
 *   relax_bdytend_core  share/module_bc.F:1221-1427   (relax_bdytend 1122-1168, _tile 1173-1219)
 *   spec_bdytend        share/module_bc.F:1430-1549
 *   spec_bdyupdate      share/module_bc.F:1955-2064
 *   zero_grad_bdy       share/module_bc.F:2219-2332
 *   flow_dep_bdy        share/module_bc.F:2335-2456
 *   spec_bdyupdate_ph   dyn_em/module_bc_em.F:17-157
 *   relax_bdy_dry       dyn_em/module_bc_em.F:161-346  (mass_weight 1716-1745)
 *   relax_bdy_scalar    dyn_em/module_bc_em.F:348-410
 *   spec_bdy_dry        dyn_em/module_bc_em.F:413-530
 *   spec_bdy_scalar     dyn_em/module_bc_em.F:658-701
 * The boundary arrays are the Fortran arrays (jms:jme, kds:kde, spec_bdy_width) of the x
 * sides and (ims:ime, kds:kde, spec_bdy_width) of the y sides (one level for mu); side
 * 0..3 = xs, xe, ys, ye.  The tile is the whole domain.
 */
#include <stdlib.h>
#include "oracle.h"

/* the zone geometry of a variable's staggering (the routines' ibs/ibe/itf/jbs/jbe/jtf/ktf);
 * var: 'u', 'v', 'm' (2-D mass), 'h' (w levels), 'w' (zero_grad_bdy's w levels), other =
 * mass levels kts..kde-1 */
typedef struct {
  int ibs, ibe, itf, jbs, jbe, jtf, ktf, twod;
} zone_t;

static zone_t zone(char var, const dims_t *d) {
  zone_t z;
  z.ibs = d->ids; z.ibe = d->ide - 1; z.itf = IMIN(d->ite, d->ide - 1);
  z.jbs = d->jds; z.jbe = d->jde - 1; z.jtf = IMIN(d->jte, d->jde - 1);
  z.ktf = d->kde - 1;
  z.twod = var == 'm';
  if (var == 'u') { z.ibe = d->ide; z.itf = IMIN(d->ite, d->ide); }
  if (var == 'v') { z.jbe = d->jde; z.jtf = IMIN(d->jte, d->jde); }
  if (var == 'm') z.ktf = d->kts;  /* the 2-D calls pass kts = kte = 1 */
  if (var == 'h' || var == 'w') z.ktf = d->kde;
  return z;
}

/* boundary array element: x sides (j, k, b), y sides (i, k, b); b is 1-based */
static inline wreal bdy(const wreal *a, int side, int ij, int k, int b, const zone_t *z, const dims_t *d) {
  const int nkb = z->twod ? 1 : d->kde - d->kds + 1;
  const int kk = z->twod ? 0 : k - d->kds;
  const size_t n = side < 2 ? (size_t)(d->jme - d->jms + 1) : (size_t)NI_;
  const int o = side < 2 ? ij - d->jms : ij - d->ims;
  return a[((size_t)(b - 1) * nkb + kk) * n + (size_t)o];
}
#define F3(a, i, k, j) (*(z.twod ? &A2(a, i, j) : &A3(a, i, k, j)))

/* The spec-zone points (those the specified rows/columns hold) in the reference's loop
 * order: y sides then, unless periodic in x, x sides.  fn(i, k, j, side, b_dist). */
#define SPEC_ZONE_LOOP(z, c, d, spec_zone, BODY)                                                         \
  do {                                                                                                   \
    const int px_ = (c)->periodic_x;                                                                     \
    if ((d)->jts - z.jbs < (spec_zone))                                                                  \
      for (int j = (d)->jts; j <= IMIN(z.jtf, z.jbs + (spec_zone) - 1); j++) {                           \
        const int b_dist = j - z.jbs, b_limit = px_ ? 0 : b_dist, side = 2;                              \
        for (int k = (d)->kts; k <= z.ktf; k++)                                                          \
          for (int i = IMAX((d)->its, b_limit + z.ibs); i <= IMIN(z.itf, z.ibe - b_limit); i++) { BODY } \
      }                                                                                                  \
    if (z.jbe - z.jtf < (spec_zone))                                                                     \
      for (int j = IMAX((d)->jts, z.jbe - (spec_zone) + 1); j <= z.jtf; j++) {                           \
        const int b_dist = z.jbe - j, b_limit = px_ ? 0 : b_dist, side = 3;                              \
        for (int k = (d)->kts; k <= z.ktf; k++)                                                          \
          for (int i = IMAX((d)->its, b_limit + z.ibs); i <= IMIN(z.itf, z.ibe - b_limit); i++) { BODY } \
      }                                                                                                  \
    if (!px_) {                                                                                          \
      if ((d)->its - z.ibs < (spec_zone))                                                                \
        for (int i = (d)->its; i <= IMIN(z.itf, z.ibs + (spec_zone) - 1); i++) {                         \
          const int b_dist = i - z.ibs, side = 0;                                                        \
          for (int k = (d)->kts; k <= z.ktf; k++)                                                        \
            for (int j = IMAX((d)->jts, b_dist + z.jbs + 1); j <= IMIN(z.jtf, z.jbe - b_dist - 1); j++) { BODY } \
        }                                                                                                \
      if (z.ibe - z.itf < (spec_zone))                                                                   \
        for (int i = IMAX((d)->its, z.ibe - (spec_zone) + 1); i <= z.itf; i++) {                         \
          const int b_dist = z.ibe - i, side = 1;                                                        \
          for (int k = (d)->kts; k <= z.ktf; k++)                                                        \
            for (int j = IMAX((d)->jts, b_dist + z.jbs + 1); j <= IMIN(z.jtf, z.jbe - b_dist - 1); j++) { BODY } \
        }                                                                                                \
    }                                                                                                    \
  } while (0)

/* share/module_bc.F:1221-1427: field_tend += fcx*fls0 - gcx*(fls1+fls2+fls3+fls4-4*fls0)
 * over the relaxation rows spec_zone..relax_zone-1 of each side */
void o_relax_bdytend(const wreal *field, wreal *tend, wreal *const *b, wreal *const *bt, char var, const cfg_t *c,
                     int spec_zone, int relax_zone, wreal dtbc, const wreal *fcx, const wreal *gcx, const dims_t *d) {
  const zone_t z = zone(var, d);
  const int px = c->periodic_x;
#define FLS(side, ij, bb, fi, fk, fj) \
  (bdy(b[side], side, ij, k, bb, &z, d) + dtbc * bdy(bt[side], side, ij, k, bb, &z, d) - F3(field, fi, fk, fj))
#define RELAX(f0, f1, f2, f3, f4, bd)                                                        \
  F3(tend, i, k, j) = F3(tend, i, k, j) + fcx[(bd)] * (f0) - gcx[(bd)] * ((f1) + (f2) + (f3) + (f4) - R(4.) * (f0))
  if (d->jts - z.jbs < relax_zone)
    for (int j = IMAX(d->jts, z.jbs + spec_zone); j <= IMIN(z.jtf, z.jbs + relax_zone - 1); j++) {
      const int b_dist = j - z.jbs, b_limit = px ? 0 : b_dist;
      for (int k = d->kts; k <= z.ktf; k++)
        for (int i = IMAX(d->its, b_limit + z.ibs); i <= IMIN(z.itf, z.ibe - b_limit); i++) {
          const int im1 = IMAX(i - 1, z.ibs), ip1 = IMIN(i + 1, z.ibe);
          const wreal f0 = FLS(2, i, b_dist + 1, i, k, j), f1 = FLS(2, im1, b_dist + 1, im1, k, j),
                      f2 = FLS(2, ip1, b_dist + 1, ip1, k, j), f3 = FLS(2, i, b_dist, i, k, j - 1),
                      f4 = FLS(2, i, b_dist + 2, i, k, j + 1);
          RELAX(f0, f1, f2, f3, f4, b_dist);
        }
    }
  if (z.jbe - z.jtf < relax_zone)
    for (int j = IMAX(d->jts, z.jbe - relax_zone + 1); j <= IMIN(z.jtf, z.jbe - spec_zone); j++) {
      const int b_dist = z.jbe - j, b_limit = px ? 0 : b_dist;
      for (int k = d->kts; k <= z.ktf; k++)
        for (int i = IMAX(d->its, b_limit + z.ibs); i <= IMIN(z.itf, z.ibe - b_limit); i++) {
          const int im1 = IMAX(i - 1, z.ibs), ip1 = IMIN(i + 1, z.ibe);
          const wreal f0 = FLS(3, i, b_dist + 1, i, k, j), f1 = FLS(3, im1, b_dist + 1, im1, k, j),
                      f2 = FLS(3, ip1, b_dist + 1, ip1, k, j), f3 = FLS(3, i, b_dist, i, k, j + 1),
                      f4 = FLS(3, i, b_dist + 2, i, k, j - 1);
          RELAX(f0, f1, f2, f3, f4, b_dist);
        }
    }
  if (!px) {
    if (d->its - z.ibs < relax_zone)
      for (int i = IMAX(d->its, z.ibs + spec_zone); i <= IMIN(z.itf, z.ibs + relax_zone - 1); i++) {
        const int b_dist = i - z.ibs;
        for (int k = d->kts; k <= z.ktf; k++)
          for (int j = IMAX(d->jts, b_dist + z.jbs + 1); j <= IMIN(z.jtf, z.jbe - b_dist - 1); j++) {
            const wreal f0 = FLS(0, j, b_dist + 1, i, k, j), f1 = FLS(0, j - 1, b_dist + 1, i, k, j - 1),
                        f2 = FLS(0, j + 1, b_dist + 1, i, k, j + 1), f3 = FLS(0, j, b_dist, i - 1, k, j),
                        f4 = FLS(0, j, b_dist + 2, i + 1, k, j);
            RELAX(f0, f1, f2, f3, f4, b_dist);
          }
      }
    if (z.ibe - z.itf < relax_zone)
      for (int i = IMAX(d->its, z.ibe - relax_zone + 1); i <= IMIN(z.itf, z.ibe - spec_zone); i++) {
        const int b_dist = z.ibe - i;
        for (int k = d->kts; k <= z.ktf; k++)
          for (int j = IMAX(d->jts, b_dist + z.jbs + 1); j <= IMIN(z.jtf, z.jbe - b_dist - 1); j++) {
            const wreal f0 = FLS(1, j, b_dist + 1, i, k, j), f1 = FLS(1, j - 1, b_dist + 1, i, k, j - 1),
                        f2 = FLS(1, j + 1, b_dist + 1, i, k, j + 1), f3 = FLS(1, j, b_dist, i + 1, k, j),
                        f4 = FLS(1, j, b_dist + 2, i - 1, k, j);
            RELAX(f0, f1, f2, f3, f4, b_dist);
          }
      }
  }
#undef FLS
#undef RELAX
}

/* share/module_bc.F:1430-1549: the spec zone's tendency is the boundary tendency */
void o_spec_bdytend(wreal *tend, wreal *const *bt, char var, const cfg_t *c, int spec_zone, const dims_t *d) {
  const zone_t z = zone(var, d);
  SPEC_ZONE_LOOP(z, c, d, spec_zone, {
    F3(tend, i, k, j) = bdy(bt[side], side, side < 2 ? j : i, k, b_dist + 1, &z, d);
  });
}

/* share/module_bc.F:1955-2064 */
void o_spec_bdyupdate(wreal *field, const wreal *tend, wreal dt, char var, const cfg_t *c, int spec_zone,
                      const dims_t *d) {
  const zone_t z = zone(var, d);
  SPEC_ZONE_LOOP(z, c, d, spec_zone, { F3(field, i, k, j) = F3(field, i, k, j) + dt * F3(tend, i, k, j); });
}

/* dyn_em/module_bc_em.F:17-157 (variable 'h') */
void o_spec_bdyupdate_ph(const wreal *ph_save, wreal *field, const wreal *tend, const wreal *mu_tend,
                         const wreal *muts, const wreal *c1, const wreal *c2, wreal dt, const cfg_t *c, int spec_zone,
                         const dims_t *d) {
  const zone_t z = zone('h', d);
  SPEC_ZONE_LOOP(z, c, d, spec_zone, {
    const wreal mu_old = A2(muts, i, j) - dt * A2(mu_tend, i, j);
    A3(field, i, k, j) = A3(field, i, k, j) * (A1(c1, k) * mu_old + A1(c2, k)) / (A1(c1, k) * A2(muts, i, j) + A1(c2, k)) +
                         dt * A3(tend, i, k, j) / (A1(c1, k) * A2(muts, i, j) + A1(c2, k)) +
                         A3(ph_save, i, k, j) * ((A1(c1, k) * mu_old + A1(c2, k)) / (A1(c1, k) * A2(muts, i, j) + A1(c2, k)) - R(1.));
  });
}

/* share/module_bc.F:2219-2332: the zone takes the values of the first row inside it */
void o_zero_grad_bdy(wreal *field, char var, const cfg_t *c, int spec_zone, const dims_t *d) {
  const zone_t z = zone(var, d);
  const int px = c->periodic_x;
  SPEC_ZONE_LOOP(z, c, d, spec_zone, {
    if (side >= 2) {
      int ii = IMIN(IMAX(i, z.ibs + spec_zone), z.ibe - spec_zone);
      if (px) ii = i;
      A3(field, i, k, j) = A3(field, ii, k, side == 2 ? z.jbs + spec_zone : z.jbe - spec_zone);
    } else {
      const int jj = IMIN(IMAX(j, z.jbs + spec_zone), z.jbe - spec_zone);
      A3(field, i, k, j) = A3(field, side == 0 ? z.ibs + spec_zone : z.ibe - spec_zone, k, jj);
    }
  });
}

/* share/module_bc.F:2335-2456: inflow points zero, outflow points the inner row's value */
void o_flow_dep_bdy(wreal *field, const wreal *u, const wreal *v, const cfg_t *c, int spec_zone, const dims_t *d) {
  const zone_t z = zone('t', d);
  const int px = c->periodic_x;
  SPEC_ZONE_LOOP(z, c, d, spec_zone, {
    if (side >= 2) {
      int ii = IMIN(IMAX(i, z.ibs + spec_zone), z.ibe - spec_zone);
      if (px) ii = i;
      const int out = side == 2 ? A3(v, i, k, j) < R(0.) : A3(v, i, k, j + 1) > R(0.);
      A3(field, i, k, j) = out ? A3(field, ii, k, side == 2 ? z.jbs + spec_zone : z.jbe - spec_zone) : R(0.);
    } else {
      const int jj = IMIN(IMAX(j, z.jbs + spec_zone), z.jbe - spec_zone);
      const int out = side == 0 ? A3(u, i, k, j) < R(0.) : A3(u, i + 1, k, j) > R(0.);
      A3(field, i, k, j) = out ? A3(field, side == 0 ? z.ibs + spec_zone : z.ibe - spec_zone, k, jj) : R(0.);
    }
  });
}

/* mass_weight (dyn_em/module_bc_em.F:1716-1745) into a full-size work array over
 * i = ids..ide-1, j = jds..jde-1 (the tile's one-point ring, clipped to the domain) and
 * levels kts..k_end */
static wreal *mass_weighted(const wreal *field, const wreal *mut, const wreal *c1, const wreal *c2, int k_end,
                            const dims_t *d) {
  const size_t n = (size_t)NI_ * NK_ * (size_t)(d->jme - d->jms + 1);
  wreal *r = (wreal *)calloc(n, sizeof(wreal));
  for (int j = IMAX(d->jts - 1, d->jds); j <= IMIN(d->jte + 1, d->jde - 1); j++)
    for (int k = d->kts; k <= k_end; k++)
      for (int i = IMAX(d->its - 1, d->ids); i <= IMIN(d->ite + 1, d->ide - 1); i++)
        A3(r, i, k, j) = A3(field, i, k, j) * (A1(c1, k) * A2(mut, i, j) + A1(c2, k));
  return r;
}

/* dyn_em/module_bc_em.F:161-346, with solve_em.F:939's arguments: the relaxation of ru, rv,
 * ph (mass weighted), t (mass weighted) and mu into u_save, v_save, ph_save, t_save and
 * mu_tend, and for a nested domain of w into w_save */
void o_relax_bdy_dry(ostate *s, const cfg_t *c, const dims_t *d) {
  const olbc *L = s->lbc;
  const int sz = c->spec_zone, rz = c->relax_zone;
  o_relax_bdytend(s->ru, s->u_save, L->u_b, L->u_bt, 'u', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
  o_relax_bdytend(s->rv, s->v_save, L->v_b, L->v_bt, 'v', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
  wreal *r = mass_weighted(s->ph_2, s->mut, s->c1f, s->c2f, d->kte, d);
  o_relax_bdytend(r, s->ph_save, L->ph_b, L->ph_bt, 'h', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
  free(r);
  r = mass_weighted(s->t_2, s->mut, s->c1h, s->c2h, d->kte - 1, d);
  o_relax_bdytend(r, s->t_save, L->t_b, L->t_bt, 't', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
  free(r);
  o_relax_bdytend(s->mu_2, s->mu_tend, L->mu_b, L->mu_bt, 'm', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
  if (c->nested) {
    r = mass_weighted(s->w_2, s->mut, s->c1f, s->c2f, d->kte, d);
    o_relax_bdytend(r, s->w_save, L->w_b, L->w_bt, 'h', c, sz, rz, L->dtbc, L->fcx, L->gcx, d);
    free(r);
  }
}

/* dyn_em/module_bc_em.F:413-530, with solve_em.F:983's arguments */
void o_spec_bdy_dry(ostate *s, const cfg_t *c, const dims_t *d) {
  const olbc *L = s->lbc;
  const int sz = c->spec_zone;
  o_spec_bdytend(s->ru_tend, L->u_bt, 'u', c, sz, d);
  o_spec_bdytend(s->rv_tend, L->v_bt, 'v', c, sz, d);
  o_spec_bdytend(s->ph_tend, L->ph_bt, 'h', c, sz, d);
  o_spec_bdytend(s->t_tend, L->t_bt, 't', c, sz, d);
  o_spec_bdytend(s->mu_tend, L->mu_bt, 'm', c, sz, d);
  if (c->nested) o_spec_bdytend(s->rw_tend, L->w_bt, 'h', c, sz, d);
}

/* dyn_em/module_bc_em.F:348-410 and 658-701 for species m (solve_em.F:2264, 2281): the
 * mass-weighted relaxation and the spec zone's boundary tendency, into q_tend */
void o_relax_spec_bdy_scalar(ostate *s, const cfg_t *c, const dims_t *d, int m) {
  const olbc *L = s->lbc;
  wreal *r = mass_weighted(s->q[m], s->mut, s->c1h, s->c2h, d->kte - 1, d);
  o_relax_bdytend(r, s->q_tend[m], L->q_b[m], L->q_bt[m], 'q', c, c->spec_zone, c->relax_zone, L->dtbc, L->fcx,
                  L->gcx, d);
  free(r);
  o_spec_bdytend(s->q_tend[m], L->q_bt[m], 'q', c, c->spec_zone, d);
}

/* whether species m takes boundary arrays (solve_em.F:2261-2262, 2809-2810): QV always,
 * other moist species with have_bcs_moist, passive scalars with have_bcs_scalar; every
 * species of a nested domain */
int o_species_has_bcs(const ostate *s, const cfg_t *c, int m) {
  if (c->nested) return 1;
  if (m == 0 && s->n_moist > 0) return 1;
  return m < s->n_moist ? c->have_bcs_moist : c->have_bcs_scalar;
}
