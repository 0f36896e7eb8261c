/*
 * Oracle: positive-definite scalar transport.  TEST INFRASTRUCTURE.
 *   advect_scalar_pd     dyn_em/module_advect_em.F:6069-7885 (h_sca_adv_order 5:
 *                        6505-6785; radiation terms of open boundaries 7255-7362; vertical
 *                        order 3: 7595-7645; limiter 7681-7779; flux divergence 7781-7883)
 *   rk_update_scalar_pd  dyn_em/module_em.F:1803-1916
 * The flux operators of advect_scalar_pd are its own statement functions (6166-6186):
 * products with the constants 7/12, 1/12, 37/60, 2/15 and 1/60, not the divisions of the
 * other advection routines, plus the first-order upwind flux of Smolarkiewicz (1989).
 */
#include <stdlib.h>
#include "oracle.h"

/* module_advect_em.F:6166-6186; ts = sign(1,time_step) as a REAL */
static inline wreal pflux4(wreal qm2, wreal qm1, wreal q0, wreal qp1) {
  return (R(7.) / R(12.)) * (q0 + qm1) - (R(1.) / R(12.)) * (qp1 + qm2);
}
static inline wreal pflux3(wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal ua, wreal ts) {
  return pflux4(qm2, qm1, q0, qp1) + ts * sgn1(ua) * (R(1.) / R(12.)) * ((qp1 - qm2) - R(3.) * (q0 - qm1));
}
static inline wreal pflux6(wreal qm3, wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal qp2) {
  return (R(37.) / R(60.)) * (q0 + qm1) - (R(2.) / R(15.)) * (qp1 + qm2) + (R(1.) / R(60.)) * (qp2 + qm3);
}
static inline wreal pflux5(wreal qm3, wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal qp2, wreal ua, wreal ts) {
  return pflux6(qm3, qm2, qm1, q0, qp1, qp2) -
         ts * sgn1(ua) * (R(1.) / R(60.)) * ((qp2 - qm3) - R(5.) * (qp1 - qm2) + R(10.) * (q0 - qm1));
}
static inline wreal rmin(wreal a, wreal b) { return a < b ? a : b; }
static inline wreal rmax(wreal a, wreal b) { return a > b ? a : b; }
static inline wreal flux_upwind(wreal qm1, wreal q0, wreal cr) {
  return R(0.5) * rmin(R(1.0), (cr + (wreal)fabs((double)cr))) * qm1 + R(0.5) * rmax(R(-1.0), (cr - (wreal)fabs((double)cr))) * q0;
}

/* 3-D temporaries dimensioned (its-1:ite+2, kts:kte, jts-1:jte+2) as in the reference */
typedef struct { wreal *p; int i0, ni, k0, nk, j0; } tmp3;
static tmp3 t3_new(const dims_t *d) {
  tmp3 t;
  t.i0 = d->its - 1; t.ni = d->ite + 2 - t.i0 + 1; t.k0 = d->kts; t.nk = d->kte - d->kts + 1; t.j0 = d->jts - 1;
  const int nj = d->jte + 2 - t.j0 + 1;
  t.p = (wreal *)calloc((size_t)t.ni * (size_t)t.nk * (size_t)nj, sizeof(wreal));
  return t;
}
#define T3(t, i, k, j) ((t).p[(((size_t)((j) - (t).j0)) * (t).nk + (size_t)((k) - (t).k0)) * (t).ni + (size_t)((i) - (t).i0)])

int o_advect_scalar_pd(const wreal *field, const wreal *field_old, wreal *tend, wreal *h_tend, wreal *z_tend,
                       const wreal *ru, const wreal *rv, const wreal *rom, const wreal *c1, const wreal *c2,
                       const wreal *mut, const wreal *mub, const wreal *mu_old, int time_step, const cfg_t *c,
                       int tenddec, const wreal *msfux, const wreal *msfuy, const wreal *msfvx, const wreal *msfvy,
                       const wreal *msftx, const wreal *msfty, const wreal *fzm, const wreal *fzp, wreal rdx, wreal rdy,
                       const wreal *rdzw, wreal dt, const dims_t *d) {
  (void)msfux; (void)msfuy; (void)msfvx; (void)msfvy;
  if (c->h_sca_adv_order != 5 || c->v_sca_adv_order != 3 || c->polar) return WRFMI_E_BADOPT;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal ts = time_step >= 0 ? R(1.) : R(-1.);
  const wreal eps = R(1.e-20);
  tmp3 fqx = t3_new(d), fqy = t3_new(d), fqz = t3_new(d), fqxl = t3_new(d), fqyl = t3_new(d), fqzl = t3_new(d);
  tmp3 flux_out = t3_new(d), ph_low = t3_new(d);
  /* 6505-6517 */
  int degrade_xs = !(c->periodic_x || c->symmetric_xs || its > ids + 3);
  int degrade_xe = !(c->periodic_x || c->symmetric_xe || ite < ide - 4);
  int degrade_ys = !(c->periodic_y || c->symmetric_ys || jts > jds + 3);
  int degrade_ye = !(c->periodic_y || c->symmetric_ye || jte < jde - 4);
  const int ktf = IMIN(kte, kde - 1);
  int i_start, i_end, j_start, j_end, i_start_f, i_end_f, j_start_f, j_end_f;
  wreal dx, dy, dz, mu, vel, cr;

  /* y flux, 6523-6585 */
  i_start = its - 1;
  i_end = IMIN(ite, ide - 1) + 1;
  j_start = jts - 1;
  j_end = IMIN(jte, jde - 1) + 1;
  j_start_f = j_start;
  j_end_f = j_end + 1;
  if (degrade_xs) i_start = IMAX(its - 1, ids);
  if (degrade_xe) i_end = IMIN(ite + 1, ide - 1);
  if (degrade_ys) { j_start = IMAX(jts - 1, jds + 1); j_start_f = jds + 3; }
  if (degrade_ye) { j_end = IMIN(jte + 1, jde - 2); j_end_f = jde - 3; }
  for (int j = j_start; j <= j_end + 1; j++) {
    const int full = j >= j_start_f && j <= j_end_f;
    const int second = !full && (j == jds + 1 || j == jde - 1);
    const int third = !full && !second && (j == jds + 2 || j == jde - 2);
    if (!full && !second && !third) continue;
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        dy = R(2.) / (A2(msftx, i, j) + A2(msftx, i, j - 1)) / rdy;
        mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j - 1) + A1(c2, k)));
        vel = A3(rv, i, k, j);
        cr = vel * dt / dy / mu;
        T3(fqyl, i, k, j) = mu * (dy / dt) * flux_upwind(A3(field_old, i, k, j - 1), A3(field_old, i, k, j), cr);
        if (full)
          T3(fqy, i, k, j) = vel * pflux5(A3(field, i, k, j - 3), A3(field, i, k, j - 2), A3(field, i, k, j - 1),
                                          A3(field, i, k, j), A3(field, i, k, j + 1), A3(field, i, k, j + 2), vel, ts);
        else if (second)
          T3(fqy, i, k, j) = R(0.5) * A3(rv, i, k, j) * (A3(field, i, k, j) + A3(field, i, k, j - 1));
        else
          T3(fqy, i, k, j) = vel * pflux3(A3(field, i, k, j - 2), A3(field, i, k, j - 1), A3(field, i, k, j),
                                          A3(field, i, k, j + 1), vel, ts);
        T3(fqy, i, k, j) = T3(fqy, i, k, j) - T3(fqyl, i, k, j);
      }
  }

  /* x flux, 6658-6781 */
  i_start = its - 1;
  i_end = IMIN(ite, ide - 1) + 1;
  i_start_f = i_start;
  i_end_f = i_end + 1;
  j_start = jts - 1;
  j_end = IMIN(jte, jde - 1) + 1;
  if (degrade_ys) j_start = IMAX(jts - 1, jds);
  if (degrade_ye) j_end = IMIN(jte + 1, jde - 1);
  if (degrade_xs) { i_start = IMAX(ids + 1, its - 1); i_start_f = ids + 3; }
  if (degrade_xe) { i_end = IMIN(ide - 2, ite + 1); i_end_f = ide - 3; }
  for (int j = j_start; j <= j_end; j++) {
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start_f; i <= i_end_f; i++) {
        dx = R(2.) / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
        mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
        vel = A3(ru, i, k, j);
        cr = vel * dt / dx / mu;
        T3(fqxl, i, k, j) = mu * (dx / dt) * flux_upwind(A3(field_old, i - 1, k, j), A3(field_old, i, k, j), cr);
        T3(fqx, i, k, j) = vel * pflux5(A3(field, i - 3, k, j), A3(field, i - 2, k, j), A3(field, i - 1, k, j),
                                        A3(field, i, k, j), A3(field, i + 1, k, j), A3(field, i + 2, k, j), vel, ts);
        T3(fqx, i, k, j) = T3(fqx, i, k, j) - T3(fqxl, i, k, j);
      }
    if (degrade_xs)
      for (int i = i_start; i <= i_start_f - 1; i++) {
        if (i == ids + 1)  /* second order; vel = ru/mu and cr = vel*dt/dx here (6733-6736) */
          for (int k = kts; k <= ktf; k++) {
            dx = R(2.) / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
            mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
            vel = A3(ru, i, k, j) / mu;
            cr = vel * dt / dx;
            T3(fqxl, i, k, j) = mu * (dx / dt) * flux_upwind(A3(field_old, i - 1, k, j), A3(field_old, i, k, j), cr);
            T3(fqx, i, k, j) = R(0.5) * (A3(ru, i, k, j)) * (A3(field, i, k, j) + A3(field, i - 1, k, j));
            T3(fqx, i, k, j) = T3(fqx, i, k, j) - T3(fqxl, i, k, j);
          }
        if (i == ids + 2)
          for (int k = kts; k <= ktf; k++) {
            dx = R(2.) / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
            mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
            vel = A3(ru, i, k, j);
            cr = vel * dt / dx / mu;
            T3(fqxl, i, k, j) = mu * (dx / dt) * flux_upwind(A3(field_old, i - 1, k, j), A3(field_old, i, k, j), cr);
            T3(fqx, i, k, j) = vel * pflux3(A3(field, i - 2, k, j), A3(field, i - 1, k, j), A3(field, i, k, j),
                                            A3(field, i + 1, k, j), vel, ts);
            T3(fqx, i, k, j) = T3(fqx, i, k, j) - T3(fqxl, i, k, j);
          }
      }
    if (degrade_xe)
      for (int i = i_end_f + 1; i <= i_end + 1; i++) {
        if (i == ide - 1)
          for (int k = kts; k <= ktf; k++) {
            dx = R(2.) / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
            mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
            vel = A3(ru, i, k, j);
            cr = vel * dt / dx / mu;
            T3(fqxl, i, k, j) = mu * (dx / dt) * flux_upwind(A3(field_old, i - 1, k, j), A3(field_old, i, k, j), cr);
            T3(fqx, i, k, j) = R(0.5) * (A3(ru, i, k, j)) * (A3(field, i, k, j) + A3(field, i - 1, k, j));
            T3(fqx, i, k, j) = T3(fqx, i, k, j) - T3(fqxl, i, k, j);
          }
        if (i == ide - 2)
          for (int k = kts; k <= ktf; k++) {
            dx = R(2.) / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
            mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
            vel = A3(ru, i, k, j);
            cr = vel * dt / dx / mu;
            T3(fqxl, i, k, j) = mu * (dx / dt) * flux_upwind(A3(field_old, i - 1, k, j), A3(field_old, i, k, j), cr);
            T3(fqx, i, k, j) = vel * pflux3(A3(field, i - 2, k, j), A3(field, i - 1, k, j), A3(field, i, k, j),
                                            A3(field, i + 1, k, j), vel, ts);
            T3(fqx, i, k, j) = T3(fqx, i, k, j) - T3(fqxl, i, k, j);
          }
      }
  }

  /* radiation terms of open boundaries, 7255-7326 */
  i_start = its;
  i_end = IMIN(ite, ide - 1);
  j_start = jts;
  j_end = IMIN(jte, jde - 1);
  if (c->open_xs && its == ids)
    for (int j = j_start; j <= j_end; j++)
      for (int k = kts; k <= ktf; k++) {
        wreal ub = rmin(R(0.5) * (A3(ru, its, k, j) + A3(ru, its + 1, k, j)), R(0.));
        A3(tend, its, k, j) = A3(tend, its, k, j) -
                              rdx * (ub * (A3(field_old, its + 1, k, j) - A3(field_old, its, k, j)) +
                                     A3(field, its, k, j) * (A3(ru, its + 1, k, j) - A3(ru, its, k, j)));
      }
  if (c->open_xe && ite == ide)
    for (int j = j_start; j <= j_end; j++)
      for (int k = kts; k <= ktf; k++) {
        wreal ub = rmax(R(0.5) * (A3(ru, ite - 1, k, j) + A3(ru, ite, k, j)), R(0.));
        A3(tend, i_end, k, j) = A3(tend, i_end, k, j) -
                                rdx * (ub * (A3(field_old, i_end, k, j) - A3(field_old, i_end - 1, k, j)) +
                                       A3(field, i_end, k, j) * (A3(ru, ite, k, j) - A3(ru, ite - 1, k, j)));
      }
  if (c->open_ys && jts == jds)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        wreal vb = rmin(R(0.5) * (A3(rv, i, k, jts) + A3(rv, i, k, jts + 1)), R(0.));
        A3(tend, i, k, jts) = A3(tend, i, k, jts) -
                              rdy * (vb * (A3(field_old, i, k, jts + 1) - A3(field_old, i, k, jts)) +
                                     A3(field, i, k, jts) * (A3(rv, i, k, jts + 1) - A3(rv, i, k, jts)));
      }
  if (c->open_ye && jte == jde)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        wreal vb = rmax(R(0.5) * (A3(rv, i, k, jte - 1) + A3(rv, i, k, jte)), R(0.));
        A3(tend, i, k, j_end) = A3(tend, i, k, j_end) -
                                rdy * (vb * (A3(field_old, i, k, j_end) - A3(field_old, i, k, j_end - 1)) +
                                       A3(field, i, k, j_end) * (A3(rv, i, k, jte) - A3(rv, i, k, jte - 1)));
      }

  /* vertical flux, order 3: 7364-7377, 7595-7645 */
  i_start = its - 1;
  i_end = IMIN(ite, ide - 1) + 1;
  j_start = jts - 1;
  j_end = IMIN(jte, jde - 1) + 1;
  if (degrade_xs) i_start = IMAX(its - 1, ids);
  if (degrade_xe) i_end = IMIN(ite + 1, ide - 1);
  if (degrade_ys) j_start = IMAX(jts - 1, jds);
  if (degrade_ye) j_end = IMIN(jte + 1, jde - 1);
  for (int j = j_start; j <= j_end; j++) {
    for (int i = i_start; i <= i_end; i++) {
      T3(fqz, i, 1, j) = R(0.);
      T3(fqzl, i, 1, j) = R(0.);
      T3(fqz, i, kde, j) = R(0.);
      T3(fqzl, i, kde, j) = R(0.);
    }
    for (int k = kts + 2; k <= ktf - 1; k++)
      for (int i = i_start; i <= i_end; i++) {
        dz = R(2.) / (A1(rdzw, k) + A1(rdzw, k - 1));
        mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j) + A1(c2, k)));
        vel = A3(rom, i, k, j);
        cr = vel * dt / dz / mu;
        T3(fqzl, i, k, j) = mu * (dz / dt) * flux_upwind(A3(field_old, i, k - 1, j), A3(field_old, i, k, j), cr);
        T3(fqz, i, k, j) = vel * pflux3(A3(field, i, k - 2, j), A3(field, i, k - 1, j), A3(field, i, k, j),
                                        A3(field, i, k + 1, j), -vel, ts);
        T3(fqz, i, k, j) = T3(fqz, i, k, j) - T3(fqzl, i, k, j);
      }
    for (int i = i_start; i <= i_end; i++) {
      for (int pass = 0; pass < 2; pass++) {
        const int k = pass == 0 ? kts + 1 : ktf;
        dz = R(2.) / (A1(rdzw, k) + A1(rdzw, k - 1));
        mu = R(0.5) * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j) + A1(c2, k)));
        vel = A3(rom, i, k, j);
        cr = vel * dt / dz / mu;
        T3(fqzl, i, k, j) = mu * (dz / dt) * flux_upwind(A3(field_old, i, k - 1, j), A3(field_old, i, k, j), cr);
        T3(fqz, i, k, j) = A3(rom, i, k, j) * (A1(fzm, k) * A3(field, i, k, j) + A1(fzp, k) * A3(field, i, k - 1, j));
        T3(fqz, i, k, j) = T3(fqz, i, k, j) - T3(fqzl, i, k, j);
      }
    }
  }

  /* positive-definite limiter, 7681-7779 */
  i_start = its - 1;
  i_end = IMIN(ite, ide - 1) + 1;
  j_start = jts - 1;
  j_end = IMIN(jte, jde - 1) + 1;
  if (degrade_xs) i_start = IMAX(its - 1, ids);
  if (degrade_xe) i_end = IMIN(ite + 1, ide - 1);
  if (degrade_ys) j_start = IMAX(jts - 1, jds);
  if (degrade_ye) j_end = IMIN(jte + 1, jde - 1);
  if (c->specified || c->nested) {
    if (degrade_xs) i_start = IMAX(its - 1, ids + 1);
    if (degrade_xe) i_end = IMIN(ite + 1, ide - 2);
    if (degrade_ys) j_start = IMAX(jts - 1, jds + 1);
    if (degrade_ye) j_end = IMIN(jte + 1, jde - 2);
  }
  if (c->open_xs && degrade_xs) i_start = IMAX(its - 1, ids + 1);
  if (c->open_xe && degrade_xe) i_end = IMIN(ite + 1, ide - 2);
  if (c->open_ys && degrade_ys) j_start = IMAX(jts - 1, jds + 1);
  if (c->open_ye && degrade_ye) j_end = IMIN(jte + 1, jde - 2);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        T3(ph_low, i, k, j) = ((A1(c1, k) * A2(mub, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mu_old, i, j))) * A3(field_old, i, k, j) -
                              dt * (A2(msftx, i, j) * A2(msfty, i, j) *
                                        (rdx * (T3(fqxl, i + 1, k, j) - T3(fqxl, i, k, j)) +
                                         rdy * (T3(fqyl, i, k, j + 1) - T3(fqyl, i, k, j))) +
                                    A2(msfty, i, j) * A1(rdzw, k) * (T3(fqzl, i, k + 1, j) - T3(fqzl, i, k, j)));
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        T3(flux_out, i, k, j) = dt * ((A2(msftx, i, j) * A2(msfty, i, j)) *
                                          (rdx * (rmax(R(0.), T3(fqx, i + 1, k, j)) - rmin(R(0.), T3(fqx, i, k, j))) +
                                           rdy * (rmax(R(0.), T3(fqy, i, k, j + 1)) - rmin(R(0.), T3(fqy, i, k, j)))) +
                                      A2(msfty, i, j) * A1(rdzw, k) *
                                          (rmin(R(0.), T3(fqz, i, k + 1, j)) - rmax(R(0.), T3(fqz, i, k, j))));
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        if (T3(flux_out, i, k, j) > T3(ph_low, i, k, j)) {
          wreal scale = rmax(R(0.), T3(ph_low, i, k, j) / (T3(flux_out, i, k, j) + eps));
          if (T3(fqx, i + 1, k, j) > R(0.)) T3(fqx, i + 1, k, j) = scale * T3(fqx, i + 1, k, j);
          if (T3(fqx, i, k, j) < R(0.)) T3(fqx, i, k, j) = scale * T3(fqx, i, k, j);
          if (T3(fqy, i, k, j + 1) > R(0.)) T3(fqy, i, k, j + 1) = scale * T3(fqy, i, k, j + 1);
          if (T3(fqy, i, k, j) < R(0.)) T3(fqy, i, k, j) = scale * T3(fqy, i, k, j);
          if (T3(fqz, i, k + 1, j) < R(0.)) T3(fqz, i, k + 1, j) = scale * T3(fqz, i, k + 1, j);
          if (T3(fqz, i, k, j) > R(0.)) T3(fqz, i, k, j) = scale * T3(fqz, i, k, j);
        }

  /* flux divergence, 7781-7883 */
  i_start = its;
  i_end = IMIN(ite, ide - 1);
  j_start = jts;
  j_end = IMIN(jte, jde - 1);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        const wreal zt = A1(rdzw, k) * (T3(fqz, i, k + 1, j) - T3(fqz, i, k, j) + T3(fqzl, i, k + 1, j) - T3(fqzl, i, k, j));
        A3(tend, i, k, j) = A3(tend, i, k, j) - zt;
        if (tenddec) A3(z_tend, i, k, j) = R(0.) - zt;
      }
  if (degrade_xs) i_start = IMAX(its, ids + 1);
  if (degrade_xe) i_end = IMIN(ite, ide - 2);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        const wreal xt = A2(msftx, i, j) * (rdx * (T3(fqx, i + 1, k, j) - T3(fqx, i, k, j) + T3(fqxl, i + 1, k, j) - T3(fqxl, i, k, j)));
        A3(tend, i, k, j) = A3(tend, i, k, j) - xt;
        if (tenddec) A3(h_tend, i, k, j) = R(0.) - xt;
      }
  i_start = its;
  i_end = IMIN(ite, ide - 1);
  if (degrade_ys) j_start = IMAX(jts, jds + 1);
  if (degrade_ye) j_end = IMIN(jte, jde - 2);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++) {
        const wreal yt = A2(msftx, i, j) * (rdy * (T3(fqy, i, k, j + 1) - T3(fqy, i, k, j) + T3(fqyl, i, k, j + 1) - T3(fqyl, i, k, j)));
        A3(tend, i, k, j) = A3(tend, i, k, j) - yt;
        if (tenddec) A3(h_tend, i, k, j) = A3(h_tend, i, k, j) - yt;
      }
  free(fqx.p); free(fqy.p); free(fqz.p); free(fqxl.p); free(fqyl.p); free(fqzl.p); free(flux_out.p); free(ph_low.p);
  return 0;
}

/* rk_update_scalar_pd, dyn_em/module_em.F:1803-1916, for species m: the physics tendency
 * sc_tend applied to scalar (the time-n values) with mu_old and mu_new, then zeroed. */
void o_rk_update_scalar_pd(wreal *scalar, wreal *sc_tend, const wreal *c1, const wreal *c2, const wreal *mu_old,
                           const wreal *mu_new, const wreal *mu_base, wreal dt, const dims_t *d) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      const wreal muold = A2(mu_old, i, j) + A2(mu_base, i, j);
      const wreal munew = A2(mu_new, i, j) + A2(mu_base, i, j);
      for (int k = kts; k <= IMIN(kte, kde - 1); k++) {
        wreal tendency = R(0.);
        tendency = tendency + A3(sc_tend, i, k, j);
        A3(sc_tend, i, k, j) = R(0.);
        A3(scalar, i, k, j) = ((A1(c1, k) * muold + A1(c2, k)) * A3(scalar, i, k, j) + dt * tendency) /
                              (A1(c1, k) * munew + A1(c2, k));
      }
    }
}
