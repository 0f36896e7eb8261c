/*
 * Oracle: the forward-step mixing and damping terms of the first RK substep.
 * TEST INFRASTRUCTURE.
 *   sixth_order_diffusion  dyn_em/module_big_step_utilities_em.F:6220-6636
 *   rk_rayleigh_damp       dyn_em/module_big_step_utilities_em.F:5836-6110
 *   horizontal_diffusion, horizontal_diffusion_3dmp, vertical_diffusion, _mp, _3dmp, _u, _v
 *                          dyn_em/module_big_step_utilities_em.F:2715-3581 (diff_opt 1)
 * (advance_w's damp_opt = 3 branch, module_small_step_em.F:1445-1458, is in small.c.)
 * SIN is evaluated in double and rounded once, as the GPU does; 2*asin(1.) is the REAL pi.
 */
#include <math.h>
#include "oracle.h"

static inline wreal rmaxw(wreal a, wreal b) { return a > b ? a : b; }
static inline wreal rminw(wreal a, wreal b) { return a < b ? a : b; }
static inline wreal rabs(wreal a) { return (wreal)fabs((double)a); }
static inline wreal rsin(wreal a) { return (wreal)sin((double)a); }

/* module_big_step_utilities_em.F:6220-6636; name 'u', 'v', 'w' or 'm' */
void o_sixth_order_diffusion(char name, const wreal *field, wreal *tend, const wreal *mut, wreal dt, const cfg_t *c,
                             const wreal *c1, const wreal *c2, int diff_6th_opt, wreal diff_6th_factor,
                             const wreal *phb, const wreal *ph, wreal rdx, wreal rdy, const wreal *msftx,
                             const wreal *msfty, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
                             const wreal *msfvy, const dims_t *d) {
  (void)ph;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal coef = diff_6th_factor * R(0.015625) / (R(2.0) * dt);
  const int ktf = IMIN(kte, kde - 1);
  const int specified = c->specified || c->nested;
  int i_start, i_end, j_start, j_end, k_start = kts, k_end = ktf;
  if (name == 'u') {
    i_start = its; i_end = ite; j_start = jts; j_end = IMIN(jde - 1, jte);
    if (c->open_xs) i_start = IMAX(its, ids + 3);
    if (c->open_ys) j_start = IMAX(jts, jds + 3);
    if (c->open_xe) i_end = IMIN(ite, ide - 3);
    if (c->open_ye) j_end = IMIN(jte, jde - 4);
    if (specified) { i_start = IMAX(its, ids + 3); i_end = IMIN(ite, ide - 3); j_start = IMAX(jts, jds + 3); j_end = IMIN(jte, jde - 4); }
  } else if (name == 'v') {
    i_start = its; i_end = IMIN(ide - 1, ite); j_start = jts; j_end = jte;
    if (c->open_xs) i_start = IMAX(its, ids + 3);
    if (c->open_ys) j_start = IMAX(jts, jds + 3);
    if (c->open_xe) i_end = IMIN(ite, ide - 4);
    if (c->open_ye) j_end = IMIN(jte, jde - 3);
    /* 6380-6385: open_xs alone restricts all four bounds */
    if (c->open_xs || specified) { i_start = IMAX(its, ids + 3); i_end = IMIN(ite, ide - 4); j_start = IMAX(jts, jds + 3); j_end = IMIN(jte, jde - 3); }
  } else {
    i_start = its; i_end = IMIN(ide - 1, ite); j_start = jts; j_end = IMIN(jde - 1, jte);
    if (c->open_xs) i_start = IMAX(its, ids + 3);
    if (c->open_ys) j_start = IMAX(jts, jds + 3);
    if (c->open_xe) i_end = IMIN(ite, ide - 4);
    if (c->open_ye) j_end = IMIN(jte, jde - 4);
    if (specified) { i_start = IMAX(its, ids + 3); i_end = IMIN(ide - 4, ite); j_start = IMAX(jts, jds + 3); j_end = IMIN(jde - 4, jte); }
    if (name == 'w') k_start = kts + 1;
  }
#define F(i, k, j) A3(field, i, k, j)
#define MU(i, j) (A1(c1, k) * A2(mut, i, j) + A1(c2, k))
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++) {
        /* x */
        wreal fx0 = (R(10.0) * (F(i, k, j) - F(i - 1, k, j)) - R(5.0) * (F(i + 1, k, j) - F(i - 2, k, j)) + (F(i + 2, k, j) - F(i - 3, k, j)));
        wreal fx1 = (R(10.0) * (F(i + 1, k, j) - F(i, k, j)) - R(5.0) * (F(i + 2, k, j) - F(i - 1, k, j)) + (F(i + 3, k, j) - F(i - 2, k, j)));
        if (diff_6th_opt == 2) {
          if (fx0 * (F(i, k, j) - F(i - 1, k, j)) <= R(0.0)) fx0 = R(0.0);
          if (fx1 * (F(i + 1, k, j) - F(i, k, j)) <= R(0.0)) fx1 = R(0.0);
        }
        wreal sd0 = R(1.), sd1 = R(1.);
        if (c->diff_6th_slopeopt >= 1) {
          const wreal dx = R(1.) / rdx;
          const wreal dzthresh = (wreal)c->diff_6th_thresh * R(9.81) * dx;
          wreal dz0, dz1;
          if (name == 'u') {
            dz0 = rmaxw(rabs(A3(phb, i, k, j) - A3(phb, i - 1, k, j)) * A2(msfux, i, j), rabs(A3(phb, i - 1, k, j) - A3(phb, i - 2, k, j)) * A2(msfux, i - 1, j));
            dz1 = rmaxw(rabs(A3(phb, i + 1, k, j) - A3(phb, i, k, j)) * A2(msfux, i + 1, j), rabs(A3(phb, i, k, j) - A3(phb, i - 1, k, j)) * A2(msfux, i, j));
          } else if (name == 'v') {
            dz0 = rmaxw(rabs(A3(phb, i, k, j) - A3(phb, i - 1, k, j)) * A2(msfux, i, j), rabs(A3(phb, i, k, j - 1) - A3(phb, i - 1, k, j - 1)) * A2(msfux, i, j - 1));
            dz1 = rmaxw(rabs(A3(phb, i + 1, k, j) - A3(phb, i, k, j)) * A2(msfux, i + 1, j), rabs(A3(phb, i + 1, k, j - 1) - A3(phb, i, k, j - 1)) * A2(msfux, i + 1, j - 1));
          } else {
            dz0 = rabs(A3(phb, i, k, j) - A3(phb, i - 1, k, j)) * A2(msfux, i, j);
            dz1 = rabs(A3(phb, i + 1, k, j) - A3(phb, i, k, j)) * A2(msfux, i + 1, j);
          }
          sd0 = rmaxw(R(1.) - (dz0 / dzthresh), R(0.));
          sd1 = rmaxw(R(1.) - (dz1 / dzthresh), R(0.));
        }
        wreal m0, m1, tx;
        if (name == 'u') {
          m0 = MU(i - 1, j);
          m1 = MU(i, j);
          tx = coef * A2(msfux, i, j) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
        } else if (name == 'v') {
          m0 = R(0.25) * (MU(i - 1, j - 1) + MU(i, j - 1) + MU(i - 1, j) + MU(i, j));
          m1 = R(0.25) * (MU(i, j - 1) + MU(i + 1, j - 1) + MU(i, j) + MU(i + 1, j));
          tx = coef * A2(msfvx, i, j) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
        } else {
          m0 = R(0.5) * (MU(i - 1, j) + MU(i, j));
          m1 = R(0.5) * (MU(i, j) + MU(i + 1, j));
          tx = coef * A2(msftx, i, j) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
        }
        /* y */
        wreal fy0 = (R(10.0) * (F(i, k, j) - F(i, k, j - 1)) - R(5.0) * (F(i, k, j + 1) - F(i, k, j - 2)) + (F(i, k, j + 2) - F(i, k, j - 3)));
        wreal fy1 = (R(10.0) * (F(i, k, j + 1) - F(i, k, j)) - R(5.0) * (F(i, k, j + 2) - F(i, k, j - 1)) + (F(i, k, j + 3) - F(i, k, j - 2)));
        if (diff_6th_opt == 2) {
          if (fy0 * (F(i, k, j) - F(i, k, j - 1)) <= R(0.0)) fy0 = R(0.0);
          if (fy1 * (F(i, k, j + 1) - F(i, k, j)) <= R(0.0)) fy1 = R(0.0);
        }
        sd0 = R(1.);
        sd1 = R(1.);
        if (c->diff_6th_slopeopt >= 1) {
          const wreal dy = R(1.) / rdy;
          const wreal dzthresh = (wreal)c->diff_6th_thresh * R(9.81) * dy;
          wreal dz0, dz1;
          if (name == 'u') {
            dz0 = rmaxw(rabs(A3(phb, i, k, j) - A3(phb, i, k, j - 1)) * A2(msfvy, i, j), rabs(A3(phb, i - 1, k, j) - A3(phb, i - 1, k, j - 1)) * A2(msfvy, i - 1, j));
            dz1 = rmaxw(rabs(A3(phb, i, k, j + 1) - A3(phb, i, k, j)) * A2(msfvy, i, j + 1), rabs(A3(phb, i - 1, k, j + 1) - A3(phb, i - 1, k, j)) * A2(msfvy, i - 1, j + 1));
          } else if (name == 'v') {
            dz0 = rmaxw(rabs(A3(phb, i, k, j) - A3(phb, i, k, j - 1)) * A2(msfvy, i, j), rabs(A3(phb, i, k, j - 1) - A3(phb, i, k, j - 2)) * A2(msfvy, i, j - 1));
            dz1 = rmaxw(rabs(A3(phb, i, k, j + 1) - A3(phb, i, k, j)) * A2(msfvy, i, j + 1), rabs(A3(phb, i, k, j) - A3(phb, i, k, j - 1)) * A2(msfvy, i, j));
          } else {
            dz0 = rabs(A3(phb, i, k, j) - A3(phb, i, k, j - 1)) * A2(msfvy, i, j);
            dz1 = rabs(A3(phb, i, k, j + 1) - A3(phb, i, k, j)) * A2(msfvy, i, j + 1);
          }
          sd0 = rmaxw(R(1.) - (dz0 / dzthresh), R(0.));
          sd1 = rmaxw(R(1.) - (dz1 / dzthresh), R(0.));
        }
        wreal ty;
        if (name == 'u') {
          m0 = R(0.25) * (MU(i - 1, j - 1) + MU(i, j - 1) + MU(i - 1, j) + MU(i, j));
          m1 = R(0.25) * (MU(i - 1, j) + MU(i, j) + MU(i - 1, j + 1) + MU(i, j + 1));
          ty = coef * A2(msfuy, i, j) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
        } else if (name == 'v') {
          m0 = MU(i, j - 1);
          m1 = MU(i, j);
          ty = coef * A2(msfvy, i, j) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
        } else {
          m0 = R(0.5) * (MU(i, j - 1) + MU(i, j));
          m1 = R(0.5) * (MU(i, j) + MU(i, j + 1));
          ty = coef * A2(msfty, i, j) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
        }
        A3(tend, i, k, j) = A3(tend, i, k, j) + tx + ty;
      }
#undef F
#undef MU
}

/* the damping layer of one column from the top down (z00 by level), the reference value of
 * the base profile `base` at each of its heights, and the damping of `fld` into `tnd` */
static void damp_column(wreal *tnd, const wreal *fld, const wreal *base, const wreal *z_base, const wreal *c1,
                        const wreal *c2, wreal mu, wreal ztop, const wreal *z00, int k1, int ktf, wreal dampcoef,
                        wreal zdamp, int i, int j, const dims_t *d) {
  const wreal pii = R(2.0) * (wreal)asin(1.0);
  for (int k = k1; k <= ktf; k++) {
    int k2 = ktf;
    while (A1(z_base, k2) > z00[k]) k2 = k2 - 1;
    wreal ref;
    if (k2 + 1 > ktf)
      ref = A1(base, k2) + (A1(base, k2) - A1(base, k2 - 1)) * (z00[k] - A1(z_base, k2)) / (A1(z_base, k2) - A1(z_base, k2 - 1));
    else
      ref = A1(base, k2) + (A1(base, k2 + 1) - A1(base, k2)) * (z00[k] - A1(z_base, k2)) / (A1(z_base, k2 + 1) - A1(z_base, k2));
    wreal dcoef = R(1.0) - rminw(R(1.0), (ztop - z00[k]) / zdamp);
    dcoef = rsin(R(0.5) * pii * dcoef);
    dcoef = dcoef * dcoef;
    A3(tnd, i, k, j) = A3(tnd, i, k, j) - (A1(c1, k) * mu + A1(c2, k)) * (dcoef * dampcoef) * (A3(fld, i, k, j) - ref);
  }
}

/* module_big_step_utilities_em.F:5836-6110, with the fields rk_tendency passes (u_2, v_2,
 * w_2, t_2, ph_2 and the base profiles of the state) */
void o_rk_rayleigh_damp(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const int ktf = IMIN(kte, kde - 1);
  const wreal dampcoef = (wreal)c->dampcoef, zdamp = (wreal)c->zdamp;
  const wreal *phb = s->phb, *ph = s->ph_2;
  wreal z00[WRFMI_ORACLE_MAXK];
  /* u */
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide); i++) {
      const wreal ztop = R(0.5) * (A3(phb, i, kde, j) + A3(phb, i - 1, kde, j) + A3(ph, i, kde, j) + A3(ph, i - 1, kde, j)) / K_G;
      int k1 = ktf;
      wreal z = ztop;
      while (z >= (ztop - zdamp)) {
        z = R(0.25) * (A3(phb, i, k1, j) + A3(phb, i, k1 + 1, j) + A3(phb, i - 1, k1, j) + A3(phb, i - 1, k1 + 1, j) +
                       A3(ph, i, k1, j) + A3(ph, i, k1 + 1, j) + A3(ph, i - 1, k1, j) + A3(ph, i - 1, k1 + 1, j)) / K_G;
        z00[k1] = z;
        k1 = k1 - 1;
      }
      k1 = k1 + 2;
      damp_column(s->ru_tendf, s->u_2, s->u_base, s->z_base, s->c1h, s->c2h, A2(s->muu, i, j), ztop, z00, k1, ktf,
                  dampcoef, zdamp, i, j, d);
    }
  /* v */
  for (int j = jts; j <= IMIN(jte, jde); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      const wreal ztop = R(0.5) * (A3(phb, i, kde, j) + A3(phb, i, kde, j - 1) + A3(ph, i, kde, j) + A3(ph, i, kde, j - 1)) / K_G;
      int k1 = ktf;
      wreal z = ztop;
      while (z >= (ztop - zdamp)) {
        z = R(0.25) * (A3(phb, i, k1, j) + A3(phb, i, k1 + 1, j) + A3(phb, i, k1, j - 1) + A3(phb, i, k1 + 1, j - 1) +
                       A3(ph, i, k1, j) + A3(ph, i, k1 + 1, j) + A3(ph, i, k1, j - 1) + A3(ph, i, k1 + 1, j - 1)) / K_G;
        z00[k1] = z;
        k1 = k1 - 1;
      }
      k1 = k1 + 2;
      damp_column(s->rv_tendf, s->v_2, s->v_base, s->z_base, s->c1h, s->c2h, A2(s->muv, i, j), ztop, z00, k1, ktf,
                  dampcoef, zdamp, i, j, d);
    }
  /* w */
  const wreal pii = R(2.0) * (wreal)asin(1.0);
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      const wreal ztop = (A3(phb, i, kde, j) + A3(ph, i, kde, j)) / K_G;
      for (int k = kts; k <= IMIN(kte, kde); k++) {
        const wreal z = (A3(phb, i, k, j) + A3(ph, i, k, j)) / K_G;
        if (z >= (ztop - zdamp)) {
          wreal dcoef = R(1.0) - rminw(R(1.0), (ztop - z) / zdamp);
          dcoef = rsin(R(0.5) * pii * dcoef);
          dcoef = dcoef * dcoef;
          A3(s->rw_tendf, i, k, j) = A3(s->rw_tendf, i, k, j) -
                                     (A1(s->c1f, k) * A2(s->mut, i, j) + A1(s->c2f, k)) * (dcoef * dampcoef) * A3(s->w_2, i, k, j);
        }
      }
    }
  /* theta */
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      const wreal ztop = (A3(phb, i, kde, j) + A3(ph, i, kde, j)) / K_G;
      int k1 = ktf;
      wreal z = ztop;
      while (z >= (ztop - zdamp)) {
        z = R(0.5) * (A3(phb, i, k1, j) + A3(phb, i, k1 + 1, j) + A3(ph, i, k1, j) + A3(ph, i, k1 + 1, j)) / K_G;
        z00[k1] = z;
        k1 = k1 - 1;
      }
      k1 = k1 + 2;
      damp_column(s->t_tendf, s->t_2, s->t_base, s->z_base, s->c1h, s->c2h, A2(s->mut, i, j), ztop, z00, k1, ktf,
                  dampcoef, zdamp, i, j, d);
    }
}

/* ---- second-order diffusion of diff_opt = 1 (module_big_step_utilities_em.F:2715-3581) ---- */

/* the i/j loop bounds of horizontal_diffusion for a field of staggering `name` (2772-2914) */
static void hdiff_bounds(char name, const cfg_t *c, const dims_t *d, int *i0, int *i1, int *j0, int *j1) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  const int specified = c->specified || c->nested;
  if (name == 'u') {
    *i0 = its; *i1 = ite; *j0 = jts; *j1 = IMIN(jte, jde - 1);
    if (c->open_xs || specified) *i0 = IMAX(ids + 1, its);
    if (c->open_xe || specified) *i1 = IMIN(ide - 1, ite);
    if (c->open_ys || specified) *j0 = IMAX(jds + 1, jts);
    if (c->open_ye || specified) *j1 = IMIN(jde - 2, jte);
    if (c->periodic_x) { *i0 = its; *i1 = ite; }
  } else if (name == 'v') {
    *i0 = its; *i1 = IMIN(ite, ide - 1); *j0 = jts; *j1 = jte;
    if (c->open_xs || specified) *i0 = IMAX(ids + 1, its);
    if (c->open_xe || specified) *i1 = IMIN(ide - 2, ite);
    if (c->open_ys || specified) *j0 = IMAX(jds + 1, jts);
    if (c->open_ye || specified) *j1 = IMIN(jde - 1, jte);
    if (c->periodic_x) { *i0 = its; *i1 = IMIN(ite, ide - 1); }
    if (c->polar) { *j0 = IMAX(jds + 1, jts); *j1 = IMIN(jde - 1, jte); }
  } else {
    *i0 = its; *i1 = IMIN(ite, ide - 1); *j0 = jts; *j1 = IMIN(jte, jde - 1);
    if (c->open_xs || specified) *i0 = IMAX(ids + 1, its);
    if (c->open_xe || specified) *i1 = IMIN(ide - 2, ite);
    if (c->open_ys || specified) *j0 = IMAX(jds + 1, jts);
    if (c->open_ye || specified) *j1 = IMIN(jde - 2, jte);
    if (c->periodic_x) { *i0 = its; *i1 = IMIN(ite, ide - 1); }
  }
}

/* module_big_step_utilities_em.F:2715-2950; name 'u', 'v', 'w' or 'm' (khdif is unused
 * there: the coefficient is the array xkmhd) */
void o_horizontal_diffusion(char name, const wreal *field, wreal *tend, const wreal *mut, const wreal *c1,
                            const wreal *c2, const cfg_t *c, const wreal *msfux, const wreal *msfuy,
                            const wreal *msfvx, const wreal *msfvx_inv, const wreal *msfvy, const wreal *msftx,
                            const wreal *msfty, wreal khdif, const wreal *xkmhd, wreal rdx, wreal rdy,
                            const dims_t *d) {
  (void)khdif;
  const int ktf = IMIN(d->kte, d->kde - 1);
  int i0, i1, j0, j1;
  hdiff_bounds(name, c, d, &i0, &i1, &j0, &j1);
  const int k0 = name == 'w' ? d->kts + 1 : d->kts;
#define F(i, k, j) A3(field, i, k, j)
#define XK(i, k, j) A3(xkmhd, i, k, j)
#define CM(i, j) (A1(c1, k) * A2(mut, i, j) + A1(c2, k))
  for (int j = j0; j <= j1; j++)
    for (int k = k0; k <= ktf; k++)
      for (int i = i0; i <= i1; i++) {
        wreal mkrdxm, mkrdxp, mrdx, mkrdym, mkrdyp, mrdy;
        if (name == 'u') {
          mkrdxm = (A2(msftx, i - 1, j) / A2(msfty, i - 1, j)) * CM(i - 1, j) * XK(i - 1, k, j) * rdx;
          mkrdxp = (A2(msftx, i, j) / A2(msfty, i, j)) * CM(i, j) * XK(i, k, j) * rdx;
          mrdx = A2(msfux, i, j) * A2(msfuy, i, j) * rdx;
          mkrdym = ((A2(msfuy, i, j) + A2(msfuy, i, j - 1)) / (A2(msfux, i, j) + A2(msfux, i, j - 1))) *
                   R(0.25) * (CM(i, j) + CM(i, j - 1) + CM(i - 1, j - 1) + CM(i - 1, j)) *
                   R(0.25) * (XK(i, k, j) + XK(i, k, j - 1) + XK(i - 1, k, j - 1) + XK(i - 1, k, j)) * rdy;
          mkrdyp = ((A2(msfuy, i, j) + A2(msfuy, i, j + 1)) / (A2(msfux, i, j) + A2(msfux, i, j + 1))) *
                   R(0.25) * (CM(i, j) + CM(i, j + 1) + CM(i - 1, j + 1) + CM(i - 1, j)) *
                   R(0.25) * (XK(i, k, j) + XK(i, k, j + 1) + XK(i - 1, k, j + 1) + XK(i - 1, k, j)) * rdy;
          mrdy = A2(msfux, i, j) * A2(msfuy, i, j) * rdy;
        } else if (name == 'v') {
          mkrdxm = ((A2(msfvx, i, j) + A2(msfvx, i - 1, j)) / (A2(msfvy, i, j) + A2(msfvy, i - 1, j))) *
                   R(0.25) * (CM(i, j) + CM(i, j - 1) + CM(i - 1, j - 1) + CM(i - 1, j)) *
                   R(0.25) * (XK(i, k, j) + XK(i, k, j - 1) + XK(i - 1, k, j - 1) + XK(i - 1, k, j)) * rdx;
          mkrdxp = ((A2(msfvx, i, j) + A2(msfvx, i + 1, j)) / (A2(msfvy, i, j) + A2(msfvy, i + 1, j))) *
                   R(0.25) * (CM(i, j) + CM(i, j - 1) + CM(i + 1, j - 1) + CM(i + 1, j)) *
                   R(0.25) * (XK(i, k, j) + XK(i, k, j - 1) + XK(i + 1, k, j - 1) + XK(i + 1, k, j)) * rdx;
          mrdx = A2(msfvx, i, j) * A2(msfvy, i, j) * rdx;
          mkrdym = (A2(msfty, i, j - 1) / A2(msftx, i, j - 1)) * XK(i, k, j - 1) * rdy;
          mkrdyp = (A2(msfty, i, j) / A2(msftx, i, j)) * XK(i, k, j) * rdy;
          mrdy = A2(msfvx, i, j) * A2(msfvy, i, j) * rdy;
        } else if (name == 'w') {
          mkrdxm = (A2(msfux, i, j) / A2(msfuy, i, j)) *
                   R(0.25) * (CM(i, j) + CM(i - 1, j) + CM(i, j) + CM(i - 1, j)) *
                   R(0.25) * (XK(i, k, j) + XK(i - 1, k, j) + XK(i, k - 1, j) + XK(i - 1, k - 1, j)) * rdx;
          mkrdxp = (A2(msfux, i + 1, j) / A2(msfuy, i + 1, j)) *
                   R(0.25) * (CM(i + 1, j) + CM(i, j) + CM(i + 1, j) + CM(i, j)) *
                   R(0.25) * (XK(i + 1, k, j) + XK(i, k, j) + XK(i + 1, k - 1, j) + XK(i, k - 1, j)) * rdx;
          mrdx = A2(msftx, i, j) * A2(msfty, i, j) * rdx;
          mkrdym = (A2(msfvy, i, j) * A2(msfvx_inv, i, j)) *
                   R(0.25) * (CM(i, j) + CM(i, j - 1) + CM(i, j) + CM(i, j - 1)) *
                   R(0.25) * (XK(i, k, j) + XK(i, k, j - 1) + XK(i, k - 1, j) + XK(i, k - 1, j - 1)) * rdy;
          mkrdyp = (A2(msfvy, i, j + 1) * A2(msfvx_inv, i, j + 1)) *
                   R(0.25) * (CM(i, j + 1) + CM(i, j) + CM(i, j + 1) + CM(i, j)) *
                   R(0.25) * (XK(i, k, j + 1) + XK(i, k, j) + XK(i, k - 1, j + 1) + XK(i, k - 1, j)) * rdy;
          mrdy = A2(msftx, i, j) * A2(msfty, i, j) * rdy;
        } else {
          mkrdxm = (A2(msfux, i, j) / A2(msfuy, i, j)) * R(0.5) * (XK(i, k, j) + XK(i - 1, k, j)) * R(0.5) *
                   (CM(i, j) + CM(i - 1, j)) * rdx;
          mkrdxp = (A2(msfux, i + 1, j) / A2(msfuy, i + 1, j)) * R(0.5) * (XK(i + 1, k, j) + XK(i, k, j)) * R(0.5) *
                   (CM(i + 1, j) + CM(i, j)) * rdx;
          mrdx = A2(msftx, i, j) * A2(msfty, i, j) * rdx;
          mkrdym = (A2(msfvy, i, j) * A2(msfvx_inv, i, j)) * R(0.5) * (XK(i, k, j) + XK(i, k, j - 1)) * R(0.5) *
                   (CM(i, j) + CM(i, j - 1)) * rdy;
          mkrdyp = (A2(msfvy, i, j + 1) * A2(msfvx_inv, i, j + 1)) * R(0.5) * (XK(i, k, j + 1) + XK(i, k, j)) *
                   R(0.5) * (CM(i, j + 1) + CM(i, j)) * rdy;
          mrdy = A2(msftx, i, j) * A2(msfty, i, j) * rdy;
        }
        A3(tend, i, k, j) = A3(tend, i, k, j) +
                            (mrdx * (mkrdxp * (F(i + 1, k, j) - F(i, k, j)) - mkrdxm * (F(i, k, j) - F(i - 1, k, j))) +
                             mrdy * (mkrdyp * (F(i, k, j + 1) - F(i, k, j)) - mkrdym * (F(i, k, j) - F(i, k, j - 1))));
      }
}

/* module_big_step_utilities_em.F:2954-3060: mass-point diffusion of field - base_3d */
void o_horizontal_diffusion_3dmp(char name, const wreal *field, wreal *tend, const wreal *mut, const wreal *c1,
                                 const wreal *c2, const cfg_t *c, const wreal *base_3d, const wreal *msfux,
                                 const wreal *msfuy, const wreal *msfvx, const wreal *msfvx_inv, const wreal *msfvy,
                                 const wreal *msftx, const wreal *msfty, wreal khdif, const wreal *xkmhd, wreal rdx,
                                 wreal rdy, const dims_t *d) {
  (void)name; (void)khdif; (void)msfvx;
  const int ktf = IMIN(d->kte, d->kde - 1);
  int i0, i1, j0, j1;
  hdiff_bounds('m', c, d, &i0, &i1, &j0, &j1);
#define B(i, k, j) A3(base_3d, i, k, j)
  for (int j = j0; j <= j1; j++)
    for (int k = d->kts; k <= ktf; k++)
      for (int i = i0; i <= i1; i++) {
        const wreal mkrdxm = (A2(msfux, i, j) / A2(msfuy, i, j)) * R(0.5) * (XK(i, k, j) + XK(i - 1, k, j)) * R(0.5) *
                             (CM(i, j) + CM(i - 1, j)) * rdx;
        const wreal mkrdxp = (A2(msfux, i + 1, j) / A2(msfuy, i + 1, j)) * R(0.5) * (XK(i + 1, k, j) + XK(i, k, j)) *
                             R(0.5) * (CM(i + 1, j) + CM(i, j)) * rdx;
        const wreal mrdx = A2(msftx, i, j) * A2(msfty, i, j) * rdx;
        const wreal mkrdym = (A2(msfvy, i, j) * A2(msfvx_inv, i, j)) * R(0.5) * (XK(i, k, j) + XK(i, k, j - 1)) *
                             R(0.5) * (CM(i, j) + CM(i, j - 1)) * rdy;
        const wreal mkrdyp = (A2(msfvy, i, j + 1) * A2(msfvx_inv, i, j + 1)) * R(0.5) *
                             (XK(i, k, j + 1) + XK(i, k, j)) * R(0.5) * (CM(i, j + 1) + CM(i, j)) * rdy;
        const wreal mrdy = A2(msftx, i, j) * A2(msfty, i, j) * rdy;
        A3(tend, i, k, j) =
            A3(tend, i, k, j) +
            (mrdx * (mkrdxp * (F(i + 1, k, j) - F(i, k, j) - B(i + 1, k, j) + B(i, k, j)) -
                     mkrdxm * (F(i, k, j) - F(i - 1, k, j) - B(i, k, j) + B(i - 1, k, j))) +
             mrdy * (mkrdyp * (F(i, k, j + 1) - F(i, k, j) - B(i, k, j + 1) + B(i, k, j)) -
                     mkrdym * (F(i, k, j) - F(i, k, j - 1) - B(i, k, j) + B(i, k, j - 1))));
      }
#undef B
#undef F
#undef XK
#undef CM
}

/* the mass-level form shared by vertical_diffusion('m'), _mp and _3dmp (3113-3376): the
 * flux of field - base between levels k and k+1 (base1 a profile, base3 a 3-D array, or
 * neither), zero-gradient below the lowest level and zero through the top */
static void vdiff_mass(const wreal *field, wreal *tend, const wreal *base1, const wreal *base3, const wreal *c1,
                       const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn, const wreal *rdnw,
                       wreal kvdif, const dims_t *d) {
  const int ktf = IMIN(d->kte, d->kde - 1), kts = d->kts;
  wreal vflux[WRFMI_ORACLE_MAXK + 2];
  for (int j = d->jts; j <= IMIN(d->jte, d->jde - 1); j++)
    for (int i = d->its; i <= IMIN(d->ite, d->ide - 1); i++) {
      for (int k = kts; k <= ktf - 1; k++) {
        wreal df = A3(field, i, k + 1, j) - A3(field, i, k, j);
        if (base1) df = df - A1(base1, k + 1) + A1(base1, k);
        if (base3) df = df - A3(base3, i, k + 1, j) + A3(base3, i, k, j);
        vflux[k] = kvdif * A1(rdn, k + 1) / (R(0.5) * (A3(alt, i, k, j) + A3(alt, i, k + 1, j))) * df;
      }
      vflux[kts - 1] = vflux[kts];
      vflux[ktf] = R(0.);
      for (int k = kts; k <= ktf; k++)
        A3(tend, i, k, j) = A3(tend, i, k, j) + K_G * K_G / (A1(c1, k) * A2(mut, i, j) + A1(c2, k)) /
                                                    A3(alt, i, k, j) * A1(rdnw, k) * (vflux[k] - vflux[k - 1]);
    }
}

/* module_big_step_utilities_em.F:3064-3187; name 'w' or 'm' */
void o_vertical_diffusion(char name, const wreal *field, wreal *tend, const cfg_t *c, const wreal *c1,
                          const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn, const wreal *rdnw,
                          wreal kvdif, const dims_t *d) {
  (void)c;
  if (name != 'w') {
    vdiff_mass(field, tend, NULL, NULL, c1, c2, alt, mut, rdn, rdnw, kvdif, d);
    return;
  }
  const int ktf = IMIN(d->kte, d->kde - 1), kts = d->kts;
  wreal vflux[WRFMI_ORACLE_MAXK + 2];
  for (int j = d->jts; j <= IMIN(d->jte, d->jde - 1); j++)
    for (int i = d->its; i <= IMIN(d->ite, d->ide - 1); i++) {
      for (int k = kts; k <= ktf - 1; k++)
        vflux[k] = (kvdif / A3(alt, i, k, j)) * A1(rdnw, k) * (A3(field, i, k + 1, j) - A3(field, i, k, j));
      vflux[ktf] = R(0.);
      for (int k = kts + 1; k <= ktf; k++)
        A3(tend, i, k, j) = A3(tend, i, k, j) + A1(rdn, k) * K_G * K_G / (A1(c1, k) * A2(mut, i, j) + A1(c2, k)) /
                                                    (R(0.5) * (A3(alt, i, k, j) + A3(alt, i, k - 1, j))) *
                                                    (vflux[k] - vflux[k - 1]);
    }
}

/* module_big_step_utilities_em.F:3192-3281 */
void o_vertical_diffusion_mp(const wreal *field, wreal *tend, const cfg_t *c, const wreal *base, const wreal *c1,
                             const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn, const wreal *rdnw,
                             wreal kvdif, const dims_t *d) {
  (void)c;
  vdiff_mass(field, tend, base, NULL, c1, c2, alt, mut, rdn, rdnw, kvdif, d);
}

/* module_big_step_utilities_em.F:3286-3376 */
void o_vertical_diffusion_3dmp(const wreal *field, wreal *tend, const cfg_t *c, const wreal *base_3d,
                               const wreal *c1, const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn,
                               const wreal *rdnw, wreal kvdif, const dims_t *d) {
  (void)c;
  vdiff_mass(field, tend, NULL, base_3d, c1, c2, alt, mut, rdn, rdnw, kvdif, d);
}

/* module_big_step_utilities_em.F:3382-3480 (uv = 'u') and 3485-3581 (uv = 'v'): the flux of
 * field - base profile at the u or v column, the tendency on levels kts..ktf-1 */
static void vdiff_uv(char uv, const wreal *field, wreal *tend, const cfg_t *c, const wreal *base, const wreal *c1h,
                     const wreal *c2h, const wreal *alt, const wreal *mu, const wreal *rdn, const wreal *rdnw,
                     wreal kvdif, const dims_t *d) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts;
  const int ktf = IMIN(d->kte, d->kde - 1);
  const int specified = c->specified || c->nested;
  int i0, i1, j0 = jts, j1 = IMIN(jte, jde - 1), di, dj;
  if (uv == 'u') {
    i0 = its; i1 = ite; di = 1; dj = 0;
    if (c->open_xs || specified) i0 = IMAX(ids + 1, its);
    if (c->open_xe || specified) i1 = IMIN(ide - 1, ite);
    if (c->periodic_x) { i0 = its; i1 = ite; }
  } else {
    i0 = its; i1 = IMIN(ite, ide - 1); di = 0; dj = 1;
    if (c->open_ys || specified) j0 = IMAX(jds + 1, jts);
    if (c->open_ye || specified) j1 = IMIN(jde - 1, jte);
  }
  wreal vflux[WRFMI_ORACLE_MAXK + 2];
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      const int im = i - di, jm = j - dj;
      for (int k = kts; k <= ktf - 1; k++)
        vflux[k] = kvdif * A1(rdn, k + 1) /
                   (R(0.25) * (A3(alt, i, k, j) + A3(alt, im, k, jm) + A3(alt, i, k + 1, j) + A3(alt, im, k + 1, jm))) *
                   (A3(field, i, k + 1, j) - A3(field, i, k, j) - A1(base, k + 1) + A1(base, k));
      vflux[kts - 1] = vflux[kts];
      vflux[ktf] = R(0.);
      for (int k = kts; k <= ktf - 1; k++)
        A3(tend, i, k, j) = A3(tend, i, k, j) + K_G * K_G * A1(rdnw, k) / (A1(c1h, k) * A2(mu, i, j) + A1(c2h, k)) /
                                                    (R(0.5) * (A3(alt, im, k, jm) + A3(alt, i, k, j))) *
                                                    (vflux[k] - vflux[k - 1]);
    }
}

void o_vertical_diffusion_u(const wreal *field, wreal *tend, const cfg_t *c, const wreal *u_base, const wreal *c1h,
                            const wreal *c2h, const wreal *alt, const wreal *muu, const wreal *rdn, const wreal *rdnw,
                            wreal kvdif, const dims_t *d) {
  vdiff_uv('u', field, tend, c, u_base, c1h, c2h, alt, muu, rdn, rdnw, kvdif, d);
}

void o_vertical_diffusion_v(const wreal *field, wreal *tend, const cfg_t *c, const wreal *v_base, const wreal *c1h,
                            const wreal *c2h, const wreal *alt, const wreal *muv, const wreal *rdn, const wreal *rdnw,
                            wreal kvdif, const dims_t *d) {
  vdiff_uv('v', field, tend, c, v_base, c1h, c2h, alt, muv, rdn, rdnw, kvdif, d);
}
