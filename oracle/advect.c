/*
 * Oracle: flux-form advection, 5th-order horizontal / 3rd-order vertical.
 * TEST INFRASTRUCTURE.
 *   advect_u       dyn_em/module_advect_em.F:126-1526   (h5 486-745, open 1242-1334, v3 1470-1498)
 *   advect_v       dyn_em/module_advect_em.F:1530-3024  (h5 1904-2172, open 2694-2835, v3 2961-2996)
 *   advect_scalar  dyn_em/module_advect_em.F:3029-4359  (h5 3394-3649, open 4104-4177, v3 4306-4333)
 *   advect_w       dyn_em/module_advect_em.F:4364-6064  (h5 4784-5096, open 5690-5846, v3 5996-6031)
 * Flux operators: module_advect_em.F:199-213.  Each routine ADDS to `tend`.
 */
#include <stdlib.h>
#include "oracle.h"

/* module_advect_em.F:199-213; ts = sign(1,time_step) as a REAL */
static inline wreal flux4(wreal qm2, wreal qm1, wreal q0, wreal qp1) {
  return (R(7.) * (q0 + qm1) - (qp1 + qm2)) / R(12.0);
}
static inline wreal flux3(wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal ua, wreal ts) {
  return flux4(qm2, qm1, q0, qp1) + ts * sgn1(ua) * ((qp1 - qm2) - R(3.) * (q0 - qm1)) / R(12.0);
}
static inline wreal flux6(wreal qm3, wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal qp2) {
  return (R(37.) * (q0 + qm1) - R(8.) * (qp1 + qm2) + (qp2 + qm3)) / R(60.0);
}
static inline wreal flux5(wreal qm3, wreal qm2, wreal qm1, wreal q0, wreal qp1, wreal qp2, wreal ua, wreal ts) {
  return flux6(qm3, qm2, qm1, q0, qp1, qp2) -
         ts * sgn1(ua) * ((qp2 - qm3) - R(5.) * (qp1 - qm2) + R(10.) * (q0 - qm1)) / R(60.0);
}

static int orders_ok(const cfg_t *c, int h, int v) {
  return h == 5 && v == 3 && !c->polar;
}

/* temporaries indexed like the reference: fqx(its-1:ite+1, kts:kte), fqy(its:ite, kts:kte, 2), vflux(its:ite, kts:kte) */
typedef struct { wreal *p; int i0, ni, k0; } tmp2;
static tmp2 tmp_new(int i0, int i1, int k0, int k1) {
  tmp2 t; t.i0 = i0; t.ni = i1 - i0 + 1; t.k0 = k0;
  t.p = (wreal *)calloc((size_t)t.ni * (size_t)(k1 - k0 + 1), sizeof(wreal));
  return t;
}
#define T2(t, i, k) ((t).p[(size_t)((k) - (t).k0) * (t).ni + (size_t)((i) - (t).i0)])

int o_advect_u(const wreal *u, const wreal *u_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzw, const dims_t *d) {
  (void)msfuy; (void)msfvx; (void)msfvy; (void)msftx; (void)msfty;
  if (!orders_ok(c, c->h_mom_adv_order, c->v_mom_adv_order)) return WRFMI_E_BADOPT;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal ts = time_step >= 0 ? R(1.) : R(-1.);
  const int specified = c->specified || c->nested;
  const int ktf = IMIN(kte, kde - 1);
  int degrade_xs = !(c->periodic_x || c->symmetric_xs || its > ids + 3);
  int degrade_xe = !(c->periodic_x || c->symmetric_xe || ite < ide - 2);
  int degrade_ys = !(c->periodic_y || c->symmetric_ys || jts > jds + 3);
  int degrade_ye = !(c->periodic_y || c->symmetric_ye || jte < jde - 4);
  tmp2 fqx = tmp_new(its - 1, ite + 1, kts, kte);
  tmp2 fqy0 = tmp_new(its, ite, kts, kte), fqy1 = tmp_new(its, ite, kts, kte);
  tmp2 vflux = tmp_new(its, ite, kts, kte);
  /* y flux divergence */
  {
    int i_start = its, i_end = ite;
    if (c->open_xs || specified) i_start = IMAX(ids + 1, its);
    if (c->open_xe || specified) i_end = IMIN(ide - 1, ite);
    if (c->periodic_x) { i_start = its; i_end = ite; }
    int j_start = jts, j_end = IMIN(jte, jde - 1);
    int j_start_f = j_start, j_end_f = j_end + 1;
    if (degrade_ys) { j_start = IMAX(jts, jds + 1); j_start_f = jds + 3; }
    if (degrade_ye) { j_end = IMIN(jte, jde - 2); j_end_f = jde - 3; }
    tmp2 *jp1 = &fqy1, *jp0 = &fqy0;
    for (int j = j_start; j <= j_end + 1; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          if (j >= j_start_f && j <= j_end_f) {
            wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
            T2(*jp1, i, k) = vel * flux5(A3(u, i, k, j - 3), A3(u, i, k, j - 2), A3(u, i, k, j - 1),
                                         A3(u, i, k, j), A3(u, i, k, j + 1), A3(u, i, k, j + 2), vel, ts);
          } else if (j == jds + 1 || j == jde - 1) {
            T2(*jp1, i, k) = R(0.25) * (A3(rv, i, k, j) + A3(rv, i - 1, k, j)) * (A3(u, i, k, j) + A3(u, i, k, j - 1));
          } else if (j == jds + 2 || j == jde - 2) {
            wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
            T2(*jp1, i, k) = vel * flux3(A3(u, i, k, j - 2), A3(u, i, k, j - 1), A3(u, i, k, j), A3(u, i, k, j + 1), vel, ts);
          }
        }
      if (j > j_start)
        for (int k = kts; k <= ktf; k++)
          for (int i = i_start; i <= i_end; i++) {
            wreal mrdy = A2(msfux, i, j - 1) * rdy;
            A3(tend, i, k, j - 1) = A3(tend, i, k, j - 1) - mrdy * (T2(*jp1, i, k) - T2(*jp0, i, k));
          }
      tmp2 *t = jp1; jp1 = jp0; jp0 = t;
    }
  }
  /* x flux divergence */
  {
    int i_start = its, i_end = ite, j_start = jts, j_end = IMIN(jte, jde - 1);
    int i_start_f = i_start, i_end_f = i_end + 1;
    if (degrade_xs) { i_start = IMAX(ids + 1, its); i_start_f = ids + 3; }
    if (degrade_xe) { i_end = IMIN(ide - 1, ite); i_end_f = ide - 2; }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start_f; i <= i_end_f; i++) {
          wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
          T2(fqx, i, k) = vel * flux5(A3(u, i - 3, k, j), A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j),
                                      A3(u, i + 1, k, j), A3(u, i + 2, k, j), vel, ts);
        }
      if (degrade_xs) {
        if (i_start == ids + 1) {
          int i = ids + 1;
          for (int k = kts; k <= ktf; k++) {
            wreal ub = A3(u, i - 1, k, j);
            if (specified && A3(u, i, k, j) < R(0.)) ub = A3(u, i, k, j);
            T2(fqx, i, k) = R(0.25) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i, k, j) + ub);
          }
        }
        int i = ids + 2;
        for (int k = kts; k <= ktf; k++) {
          wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
          T2(fqx, i, k) = vel * flux3(A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), vel, ts);
        }
      }
      if (degrade_xe) {
        if (i_end == ide - 1) {
          int i = ide;
          for (int k = kts; k <= ktf; k++) {
            wreal ub = A3(u, i, k, j);
            if (specified && A3(u, i - 1, k, j) > R(0.)) ub = A3(u, i - 1, k, j);
            T2(fqx, i, k) = R(0.25) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i - 1, k, j) + ub);
          }
        }
        for (int k = kts; k <= ktf; k++) {
          int i = ide - 1;
          wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
          T2(fqx, i, k) = vel * flux3(A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), vel, ts);
        }
      }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal mrdx = A2(msfux, i, j) * rdx;
          A3(tend, i, k, j) = A3(tend, i, k, j) - mrdx * (T2(fqx, i + 1, k) - T2(fqx, i, k));
        }
    }
  }
  /* radiative lateral boundary condition in x for u */
  if (c->open_xs && its == ids)
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts; k <= ktf; k++) {
        wreal ub = A3(ru, its, k, j) - K_CB * (A1(c1, k) * A2(mut, its, j) + A1(c2, k));
        if (ub > R(0.)) ub = R(0.);
        A3(tend, its, k, j) = A3(tend, its, k, j) - rdx * ub * (A3(u_old, its + 1, k, j) - A3(u_old, its, k, j));
      }
  if (c->open_xe && ite == ide)
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts; k <= ktf; k++) {
        wreal ub = A3(ru, ite, k, j) + K_CB * (A1(c1, k) * A2(mut, ite - 1, j) + A1(c2, k));
        if (ub < R(0.)) ub = R(0.);
        A3(tend, ite, k, j) = A3(tend, ite, k, j) - rdx * ub * (A3(u_old, ite, k, j) - A3(u_old, ite - 1, k, j));
      }
  {
    int i_start = its, i_end = IMIN(ite, ide), imin = ids, imax = ide - 1;
    if (c->open_xs) { i_start = IMAX(ids + 1, its); imin = ids; }
    if (c->open_xe) { i_end = IMIN(ite, ide - 1); imax = ide - 1; }
    if (c->open_ys && jts == jds)
      for (int i = i_start; i <= i_end; i++) {
        wreal mrdy = A2(msfux, i, jts) * rdy;
        int ip = IMIN(imax, i), im = IMAX(imin, i - 1);
        for (int k = kts; k <= ktf; k++) {
          wreal vw = R(0.5) * (A3(rv, ip, k, jts) + A3(rv, im, k, jts));
          wreal vb = vw < R(0.) ? vw : R(0.);
          wreal dvm = A3(rv, ip, k, jts + 1) - A3(rv, ip, k, jts);
          wreal dvp = A3(rv, im, k, jts + 1) - A3(rv, im, k, jts);
          A3(tend, i, k, jts) = A3(tend, i, k, jts) -
                                mrdy * (vb * (A3(u_old, i, k, jts + 1) - A3(u_old, i, k, jts)) + R(0.5) * A3(u, i, k, jts) * (dvm + dvp));
        }
      }
    if (c->open_ye && jte == jde)
      for (int i = i_start; i <= i_end; i++) {
        wreal mrdy = A2(msfux, i, jte - 1) * rdy;
        int ip = IMIN(imax, i), im = IMAX(imin, i - 1);
        for (int k = kts; k <= ktf; k++) {
          wreal vw = R(0.5) * (A3(rv, ip, k, jte) + A3(rv, im, k, jte));
          wreal vb = vw > R(0.) ? vw : R(0.);
          wreal dvm = A3(rv, ip, k, jte) - A3(rv, ip, k, jte - 1);
          wreal dvp = A3(rv, im, k, jte) - A3(rv, im, k, jte - 1);
          A3(tend, i, k, jte - 1) = A3(tend, i, k, jte - 1) -
                                    mrdy * (vb * (A3(u_old, i, k, jte - 1) - A3(u_old, i, k, jte - 2)) + R(0.5) * A3(u, i, k, jte - 1) * (dvm + dvp));
        }
      }
  }
  /* vertical advection, order 3 */
  {
    int i_start = its, i_end = ite, j_start = jts, j_end = IMIN(jte, jde - 1);
    if (c->open_ys || specified) i_start = IMAX(ids + 1, its);
    if (c->open_ye || specified) i_end = IMIN(ide - 1, ite);
    if (c->periodic_x) { i_start = its; i_end = ite; }
    for (int i = i_start; i <= i_end; i++) { T2(vflux, i, kts) = R(0.); T2(vflux, i, kte) = R(0.); }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts + 2; k <= ktf - 1; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal vel = R(0.5) * (A3(rom, i - 1, k, j) + A3(rom, i, k, j));
          T2(vflux, i, k) = vel * flux3(A3(u, i, k - 2, j), A3(u, i, k - 1, j), A3(u, i, k, j), A3(u, i, k + 1, j), -vel, ts);
        }
      for (int i = i_start; i <= i_end; i++) {
        int k = kts + 1;
        T2(vflux, i, k) = R(0.5) * (A3(rom, i, k, j) + A3(rom, i - 1, k, j)) * (A1(fzm, k) * A3(u, i, k, j) + A1(fzp, k) * A3(u, i, k - 1, j));
        k = ktf;
        T2(vflux, i, k) = R(0.5) * (A3(rom, i, k, j) + A3(rom, i - 1, k, j)) * (A1(fzm, k) * A3(u, i, k, j) + A1(fzp, k) * A3(u, i, k - 1, j));
      }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(tend, i, k, j) = A3(tend, i, k, j) - A1(rdzw, k) * (T2(vflux, i, k + 1) - T2(vflux, i, k));
    }
  }
  free(fqx.p); free(fqy0.p); free(fqy1.p); free(vflux.p);
  return 0;
}

int o_advect_v(const wreal *v, const wreal *v_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzw, const dims_t *d) {
  (void)msfux; (void)msfuy; (void)msftx; (void)msfty;
  if (!orders_ok(c, c->h_mom_adv_order, c->v_mom_adv_order)) return WRFMI_E_BADOPT;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal ts = time_step >= 0 ? R(1.) : R(-1.);
  const int specified = c->specified || c->nested;
  const int ktf = IMIN(kte, kde - 1);
  int degrade_xs = !(c->periodic_x || c->symmetric_xs || its > ids + 3);
  int degrade_xe = !(c->periodic_x || c->symmetric_xe || ite < ide - 3);
  int degrade_ys = !(c->periodic_y || c->symmetric_ys || jts > jds + 3);
  int degrade_ye = !(c->periodic_y || c->symmetric_ye || jte < jde - 3);
  tmp2 fqx = tmp_new(its - 1, ite + 1, kts, kte);
  tmp2 fqy0 = tmp_new(its, ite, kts, kte), fqy1 = tmp_new(its, ite, kts, kte);
  tmp2 vflux = tmp_new(its, ite, kts, kte);
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = jte;
    int j_start_f = j_start, j_end_f = j_end + 1;
    if (degrade_ys) { j_start = IMAX(jts, jds + 1); j_start_f = jds + 3; }
    if (degrade_ye) { j_end = IMIN(jte, jde - 1); j_end_f = jde - 2; }
    tmp2 *jp1 = &fqy1, *jp0 = &fqy0;
    for (int j = j_start; j <= j_end + 1; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          if (j >= j_start_f && j <= j_end_f) {
            wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
            T2(*jp1, i, k) = vel * flux5(A3(v, i, k, j - 3), A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j),
                                         A3(v, i, k, j + 1), A3(v, i, k, j + 2), vel, ts);
          } else if (j == jds + 1) {
            wreal vb = A3(v, i, k, j - 1);
            if (specified && A3(v, i, k, j) < R(0.)) vb = A3(v, i, k, j);
            T2(*jp1, i, k) = R(0.25) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (A3(v, i, k, j) + vb);
          } else if (j == jds + 2) {
            wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
            T2(*jp1, i, k) = vel * flux3(A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), vel, ts);
          } else if (j == jde) {
            wreal vb = A3(v, i, k, j);
            if (specified && A3(v, i, k, j - 1) > R(0.)) vb = A3(v, i, k, j - 1);
            T2(*jp1, i, k) = R(0.25) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (vb + A3(v, i, k, j - 1));
          } else if (j == jde - 1) {
            wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
            T2(*jp1, i, k) = vel * flux3(A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), vel, ts);
          }
        }
      if (j > j_start)
        for (int k = kts; k <= ktf; k++)
          for (int i = i_start; i <= i_end; i++) {
            wreal mrdy = A2(msfvy, i, j - 1) * rdy;
            A3(tend, i, k, j - 1) = A3(tend, i, k, j - 1) - mrdy * (T2(*jp1, i, k) - T2(*jp0, i, k));
          }
      tmp2 *t = jp1; jp1 = jp0; jp0 = t;
    }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = jte;
    if (c->open_ys || specified) j_start = IMAX(jds + 1, jts);
    if (c->open_ye || specified) j_end = IMIN(jde - 1, jte);
    int i_start_f = i_start, i_end_f = i_end + 1;
    if (degrade_xs) { i_start = IMAX(ids + 1, its); i_start_f = IMIN(i_start + 2, ids + 3); }
    if (degrade_xe) { i_end = IMIN(ide - 2, ite); i_end_f = ide - 3; }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start_f; i <= i_end_f; i++) {
          wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
          T2(fqx, i, k) = vel * flux5(A3(v, i - 3, k, j), A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j),
                                      A3(v, i + 1, k, j), A3(v, i + 2, k, j), vel, ts);
        }
      if (degrade_xs)
        for (int i = i_start; i <= i_start_f - 1; i++) {
          if (i == ids + 1)
            for (int k = kts; k <= ktf; k++)
              T2(fqx, i, k) = R(0.25) * (A3(ru, i, k, j) + A3(ru, i, k, j - 1)) * (A3(v, i, k, j) + A3(v, i - 1, k, j));
          if (i == ids + 2)
            for (int k = kts; k <= ktf; k++) {
              wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
              T2(fqx, i, k) = vel * flux3(A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), vel, ts);
            }
        }
      if (degrade_xe)
        for (int i = i_end_f + 1; i <= i_end + 1; i++) {
          if (i == ide - 1)
            for (int k = kts; k <= ktf; k++)
              T2(fqx, i, k) = R(0.25) * (A3(ru, i_end + 1, k, j) + A3(ru, i_end + 1, k, j - 1)) *
                              (A3(v, i_end + 1, k, j) + A3(v, i_end, k, j));
          if (i == ide - 2)
            for (int k = kts; k <= ktf; k++) {
              wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
              T2(fqx, i, k) = vel * flux3(A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), vel, ts);
            }
        }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal mrdx = A2(msfvy, i, j) * rdx;
          A3(tend, i, k, j) = A3(tend, i, k, j) - mrdx * (T2(fqx, i + 1, k) - T2(fqx, i, k));
        }
    }
  }
  if (c->open_ys && jts == jds)
    for (int i = its; i <= IMIN(ite, ide - 1); i++)
      for (int k = kts; k <= ktf; k++) {
        wreal vb = A3(rv, i, k, jts) - K_CB * (A1(c1, k) * A2(mut, i, jts) + A1(c2, k));
        if (vb > R(0.)) vb = R(0.);
        A3(tend, i, k, jts) = A3(tend, i, k, jts) - rdy * vb * (A3(v_old, i, k, jts + 1) - A3(v_old, i, k, jts));
      }
  if (c->open_ye && jte == jde)
    for (int i = its; i <= IMIN(ite, ide - 1); i++)
      for (int k = kts; k <= ktf; k++) {
        wreal vb = A3(rv, i, k, jte) + K_CB * (A1(c1, k) * A2(mut, i, jte - 1) + A1(c2, k));
        if (vb < R(0.)) vb = R(0.);
        A3(tend, i, k, jte) = A3(tend, i, k, jte) - rdy * vb * (A3(v_old, i, k, jte) - A3(v_old, i, k, jte - 1));
      }
  {
    int j_start = jts, j_end = IMIN(jte, jde), jmin = jds, jmax = jde - 1;
    if (c->open_ys) { j_start = IMAX(jds + 1, jts); jmin = jds; }
    if (c->open_ye) { j_end = IMIN(jte, jde - 1); jmax = jde - 1; }
    if (c->open_xs && its == ids)
      for (int j = j_start; j <= j_end; j++) {
        wreal mrdx = A2(msfvy, its, j) * rdx;
        int jp = IMIN(jmax, j), jm = IMAX(jmin, j - 1);
        for (int k = kts; k <= ktf; k++) {
          wreal uw = R(0.5) * (A3(ru, its, k, jp) + A3(ru, its, k, jm));
          wreal ub = uw < R(0.) ? uw : R(0.);
          wreal dup = A3(ru, its + 1, k, jp) - A3(ru, its, k, jp);
          wreal dum = A3(ru, its + 1, k, jm) - A3(ru, its, k, jm);
          A3(tend, its, k, j) = A3(tend, its, k, j) -
                                mrdx * (ub * (A3(v_old, its + 1, k, j) - A3(v_old, its, k, j)) + R(0.5) * A3(v, its, k, j) * (dup + dum));
        }
      }
    if (c->open_xe && ite == ide)
      for (int j = j_start; j <= j_end; j++) {
        wreal mrdx = A2(msfvy, ite - 1, j) * rdx;
        int jp = IMIN(jmax, j), jm = IMAX(jmin, j - 1);
        for (int k = kts; k <= ktf; k++) {
          wreal uw = R(0.5) * (A3(ru, ite, k, jp) + A3(ru, ite, k, jm));
          wreal ub = uw > R(0.) ? uw : R(0.);
          wreal dup = A3(ru, ite, k, jp) - A3(ru, ite - 1, k, jp);
          wreal dum = A3(ru, ite, k, jm) - A3(ru, ite - 1, k, jm);
          A3(tend, ite - 1, k, j) = A3(tend, ite - 1, k, j) -
                                    mrdx * (ub * (A3(v_old, ite - 1, k, j) - A3(v_old, ite - 2, k, j)) + R(0.5) * A3(v, ite - 1, k, j) * (dup + dum));
        }
      }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = jte;
    for (int i = i_start; i <= i_end; i++) { T2(vflux, i, kts) = R(0.); T2(vflux, i, kte) = R(0.); }
    if (c->open_ys || specified) j_start = IMAX(jds + 1, jts);
    if (c->open_ye || specified) j_end = IMIN(jde - 1, jte);
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts + 2; k <= ktf - 1; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal vel = R(0.5) * (A3(rom, i, k, j) + A3(rom, i, k, j - 1));
          T2(vflux, i, k) = vel * flux3(A3(v, i, k - 2, j), A3(v, i, k - 1, j), A3(v, i, k, j), A3(v, i, k + 1, j), -vel, ts);
        }
      for (int i = i_start; i <= i_end; i++) {
        int k = kts + 1;
        T2(vflux, i, k) = R(0.5) * (A3(rom, i, k, j) + A3(rom, i, k, j - 1)) * (A1(fzm, k) * A3(v, i, k, j) + A1(fzp, k) * A3(v, i, k - 1, j));
        k = ktf;
        T2(vflux, i, k) = R(0.5) * (A3(rom, i, k, j) + A3(rom, i, k, j - 1)) * (A1(fzm, k) * A3(v, i, k, j) + A1(fzp, k) * A3(v, i, k - 1, j));
      }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(tend, i, k, j) = A3(tend, i, k, j) - (A2(msfvy, i, j) / A2(msfvx, i, j)) * A1(rdzw, k) * (T2(vflux, i, k + 1) - T2(vflux, i, k));
    }
  }
  free(fqx.p); free(fqy0.p); free(fqy1.p); free(vflux.p);
  return 0;
}

int o_advect_scalar(const wreal *field, const wreal *field_old, wreal *tend, const wreal *ru,
                    const wreal *rv, const wreal *rom, const wreal *c1, const wreal *c2,
                    const wreal *mut, int time_step, const cfg_t *c, const wreal *msfux,
                    const wreal *msfuy, const wreal *msfvx, const wreal *msfvy, const wreal *msftx,
                    const wreal *msfty, const wreal *fzm, const wreal *fzp, wreal rdx, wreal rdy,
                    const wreal *rdzw, const dims_t *d) {
  (void)c1; (void)c2; (void)mut; (void)msfux; (void)msfuy; (void)msfvx; (void)msfvy; (void)msfty;
  if (!orders_ok(c, c->h_sca_adv_order, c->v_sca_adv_order)) return WRFMI_E_BADOPT;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal ts = time_step >= 0 ? R(1.) : R(-1.);
  const int ktf = IMIN(kte, kde - 1);
  int degrade_xs = !(c->periodic_x || c->symmetric_xs || its > ids + 3);
  int degrade_xe = !(c->periodic_x || c->symmetric_xe || ite < ide - 3);
  int degrade_ys = !(c->periodic_y || c->symmetric_ys || jts > jds + 3);
  int degrade_ye = !(c->periodic_y || c->symmetric_ye || jte < jde - 4);
  tmp2 fqx = tmp_new(its - 1, ite + 1, kts, kte);
  tmp2 fqy0 = tmp_new(its, ite, kts, kte), fqy1 = tmp_new(its, ite, kts, kte);
  tmp2 vflux = tmp_new(its, ite, kts, kte);
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    int j_start_f = j_start, j_end_f = j_end + 1;
    if (degrade_ys) { j_start = IMAX(jts, jds + 1); j_start_f = jds + 3; }
    if (degrade_ye) { j_end = IMIN(jte, jde - 2); j_end_f = jde - 3; }
    tmp2 *jp1 = &fqy1, *jp0 = &fqy0;
    for (int j = j_start; j <= j_end + 1; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          if (j >= j_start_f && j <= j_end_f) {
            wreal vel = A3(rv, i, k, j);
            T2(*jp1, i, k) = vel * flux5(A3(field, i, k, j - 3), A3(field, i, k, j - 2), A3(field, i, k, j - 1),
                                         A3(field, i, k, j), A3(field, i, k, j + 1), A3(field, i, k, j + 2), vel, ts);
          } else if (j == jds + 1 || j == jde - 1) {
            T2(*jp1, i, k) = R(0.5) * A3(rv, i, k, j) * (A3(field, i, k, j) + A3(field, i, k, j - 1));
          } else if (j == jds + 2 || j == jde - 2) {
            wreal vel = A3(rv, i, k, j);
            T2(*jp1, i, k) = vel * flux3(A3(field, i, k, j - 2), A3(field, i, k, j - 1), A3(field, i, k, j), A3(field, i, k, j + 1), vel, ts);
          }
        }
      if (j > j_start)
        for (int k = kts; k <= ktf; k++)
          for (int i = i_start; i <= i_end; i++) {
            wreal mrdy = A2(msftx, i, j - 1) * rdy;
            A3(tend, i, k, j - 1) = A3(tend, i, k, j - 1) - mrdy * (T2(*jp1, i, k) - T2(*jp0, i, k));
          }
      tmp2 *t = jp1; jp1 = jp0; jp0 = t;
    }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    int i_start_f = i_start, i_end_f = i_end + 1;
    if (degrade_xs) { i_start = IMAX(ids + 1, its); i_start_f = IMIN(i_start + 2, ids + 3); }
    if (degrade_xe) { i_end = IMIN(ide - 2, ite); i_end_f = ide - 3; }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start_f; i <= i_end_f; i++) {
          wreal vel = A3(ru, i, k, j);
          T2(fqx, i, k) = vel * flux5(A3(field, i - 3, k, j), A3(field, i - 2, k, j), A3(field, i - 1, k, j),
                                      A3(field, i, k, j), A3(field, i + 1, k, j), A3(field, i + 2, k, j), vel, ts);
        }
      if (degrade_xs)
        for (int i = i_start; i <= i_start_f - 1; i++) {
          if (i == ids + 1)
            for (int k = kts; k <= ktf; k++)
              T2(fqx, i, k) = R(0.5) * (A3(ru, i, k, j)) * (A3(field, i, k, j) + A3(field, i - 1, k, j));
          if (i == ids + 2)
            for (int k = kts; k <= ktf; k++) {
              wreal vel = A3(ru, i, k, j);
              T2(fqx, i, k) = vel * flux3(A3(field, i - 2, k, j), A3(field, i - 1, k, j), A3(field, i, k, j), A3(field, i + 1, k, j), vel, ts);
            }
        }
      if (degrade_xe)
        for (int i = i_end_f + 1; i <= i_end + 1; i++) {
          if (i == ide - 1)
            for (int k = kts; k <= ktf; k++)
              T2(fqx, i, k) = R(0.5) * (A3(ru, i, k, j)) * (A3(field, i, k, j) + A3(field, i - 1, k, j));
          if (i == ide - 2)
            for (int k = kts; k <= ktf; k++) {
              wreal vel = A3(ru, i, k, j);
              T2(fqx, i, k) = vel * flux3(A3(field, i - 2, k, j), A3(field, i - 1, k, j), A3(field, i, k, j), A3(field, i + 1, k, j), vel, ts);
            }
        }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal mrdx = A2(msftx, i, j) * rdx;
          A3(tend, i, k, j) = A3(tend, i, k, j) - mrdx * (T2(fqx, i + 1, k) - T2(fqx, i, k));
        }
    }
  }
  {
    int i_end = IMIN(ite, ide - 1), j_end = IMIN(jte, jde - 1);
    if (c->open_xs && its == ids)
      for (int j = jts; j <= j_end; j++)
        for (int k = kts; k <= ktf; k++) {
          wreal ub = R(0.5) * (A3(ru, its, k, j) + A3(ru, its + 1, k, j));
          if (ub > R(0.)) ub = R(0.);
          A3(tend, its, k, j) = A3(tend, its, k, j) -
                                rdx * (ub * (A3(field_old, its + 1, k, j) - A3(field_old, its, k, j)) +
                                       A3(field, its, k, j) * (A3(ru, its + 1, k, j) - A3(ru, its, k, j)));
        }
    if (c->open_xe && ite == ide)
      for (int j = jts; j <= j_end; j++)
        for (int k = kts; k <= ktf; k++) {
          wreal ub = R(0.5) * (A3(ru, ite - 1, k, j) + A3(ru, ite, k, j));
          if (ub < R(0.)) ub = R(0.);
          A3(tend, i_end, k, j) = A3(tend, i_end, k, j) -
                                  rdx * (ub * (A3(field_old, i_end, k, j) - A3(field_old, i_end - 1, k, j)) +
                                         A3(field, i_end, k, j) * (A3(ru, ite, k, j) - A3(ru, ite - 1, k, j)));
        }
    if (c->open_ys && jts == jds)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= i_end; i++) {
          wreal vb = R(0.5) * (A3(rv, i, k, jts) + A3(rv, i, k, jts + 1));
          if (vb > R(0.)) vb = R(0.);
          A3(tend, i, k, jts) = A3(tend, i, k, jts) -
                                rdy * (vb * (A3(field_old, i, k, jts + 1) - A3(field_old, i, k, jts)) +
                                       A3(field, i, k, jts) * (A3(rv, i, k, jts + 1) - A3(rv, i, k, jts)));
        }
    if (c->open_ye && jte == jde)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= i_end; i++) {
          wreal vb = R(0.5) * (A3(rv, i, k, jte - 1) + A3(rv, i, k, jte));
          if (vb < R(0.)) vb = R(0.);
          A3(tend, i, k, j_end) = A3(tend, i, k, j_end) -
                                  rdy * (vb * (A3(field_old, i, k, j_end) - A3(field_old, i, k, j_end - 1)) +
                                         A3(field, i, k, j_end) * (A3(rv, i, k, jte) - A3(rv, i, k, jte - 1)));
        }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    for (int i = i_start; i <= i_end; i++) { T2(vflux, i, kts) = R(0.); T2(vflux, i, kte) = R(0.); }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts + 2; k <= ktf - 1; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal vel = A3(rom, i, k, j);
          T2(vflux, i, k) = vel * flux3(A3(field, i, k - 2, j), A3(field, i, k - 1, j), A3(field, i, k, j), A3(field, i, k + 1, j), -vel, ts);
        }
      for (int i = i_start; i <= i_end; i++) {
        int k = kts + 1;
        T2(vflux, i, k) = A3(rom, i, k, j) * (A1(fzm, k) * A3(field, i, k, j) + A1(fzp, k) * A3(field, i, k - 1, j));
        k = ktf;
        T2(vflux, i, k) = A3(rom, i, k, j) * (A1(fzm, k) * A3(field, i, k, j) + A1(fzp, k) * A3(field, i, k - 1, j));
      }
      for (int k = kts; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(tend, i, k, j) = A3(tend, i, k, j) - A1(rdzw, k) * (T2(vflux, i, k + 1) - T2(vflux, i, k));
    }
  }
  free(fqx.p); free(fqy0.p); free(fqy1.p); free(vflux.p);
  return 0;
}

int o_advect_w(const wreal *w, const wreal *w_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzu, const dims_t *d) {
  (void)c1; (void)c2; (void)mut; (void)msfux; (void)msfuy; (void)msfvx; (void)msfvy; (void)msfty;
  if (!orders_ok(c, c->h_mom_adv_order, c->v_mom_adv_order)) return WRFMI_E_BADOPT;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const wreal ts = time_step >= 0 ? R(1.) : R(-1.);
  const int ktf = IMIN(kte, kde - 1);
  int degrade_xs = !(c->periodic_x || c->symmetric_xs || its > ids + 3);
  int degrade_xe = !(c->periodic_x || c->symmetric_xe || ite < ide - 3);
  int degrade_ys = !(c->periodic_y || c->symmetric_ys || jts > jds + 3);
  int degrade_ye = !(c->periodic_y || c->symmetric_ye || jte < jde - 4);
  tmp2 fqx = tmp_new(its - 1, ite + 1, kts, kte);
  tmp2 fqy0 = tmp_new(its, ite, kts, kte), fqy1 = tmp_new(its, ite, kts, kte);
  tmp2 vflux = tmp_new(its, ite, kts, kte);
/* face velocity at w level k (k <= ktf) or at the lid (k = ktf+1) from a u/v-staggered mass flux */
#define WVEL(rr, i, k, j) ((k) <= ktf ? A1(fzm, k) * A3(rr, i, k, j) + A1(fzp, k) * A3(rr, i, (k) - 1, j) \
                                      : (R(2.) - A1(fzm, (k) - 1)) * A3(rr, i, (k) - 1, j) - A1(fzp, (k) - 1) * A3(rr, i, (k) - 2, j))
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    int j_start_f = j_start, j_end_f = j_end + 1;
    if (degrade_ys) { j_start = IMAX(jts, jds + 1); j_start_f = jds + 3; }
    if (degrade_ye) { j_end = IMIN(jte, jde - 2); j_end_f = jde - 3; }
    tmp2 *jp1 = &fqy1, *jp0 = &fqy0;
    for (int j = j_start; j <= j_end + 1; j++) {
      for (int k = kts + 1; k <= ktf + 1; k++)
        for (int i = i_start; i <= i_end; i++) {
          if (j >= j_start_f && j <= j_end_f) {
            wreal vel = WVEL(rv, i, k, j);
            T2(*jp1, i, k) = vel * flux5(A3(w, i, k, j - 3), A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j),
                                         A3(w, i, k, j + 1), A3(w, i, k, j + 2), vel, ts);
          } else if (j == jds + 1 || j == jde - 1) {
            T2(*jp1, i, k) = R(0.5) * (WVEL(rv, i, k, j)) * (A3(w, i, k, j) + A3(w, i, k, j - 1));
          } else if (j == jds + 2 || j == jde - 2) {
            wreal vel = WVEL(rv, i, k, j);
            T2(*jp1, i, k) = vel * flux3(A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j), A3(w, i, k, j + 1), vel, ts);
          }
        }
      if (j > j_start)
        for (int k = kts + 1; k <= ktf + 1; k++)
          for (int i = i_start; i <= i_end; i++) {
            wreal mrdy = A2(msftx, i, j - 1) * rdy;
            A3(tend, i, k, j - 1) = A3(tend, i, k, j - 1) - mrdy * (T2(*jp1, i, k) - T2(*jp0, i, k));
          }
      tmp2 *t = jp1; jp1 = jp0; jp0 = t;
    }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    int i_start_f = i_start, i_end_f = i_end + 1;
    if (degrade_xs) { i_start = IMAX(ids + 1, its); i_start_f = IMIN(i_start + 2, ids + 3); }
    if (degrade_xe) { i_end = IMIN(ide - 2, ite); i_end_f = ide - 3; }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts + 1; k <= ktf + 1; k++)
        for (int i = i_start_f; i <= i_end_f; i++) {
          wreal vel = WVEL(ru, i, k, j);
          T2(fqx, i, k) = vel * flux5(A3(w, i - 3, k, j), A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j),
                                      A3(w, i + 1, k, j), A3(w, i + 2, k, j), vel, ts);
        }
      if (degrade_xs)
        for (int i = i_start; i <= i_start_f - 1; i++) {
          if (i == ids + 1)
            for (int k = kts + 1; k <= ktf + 1; k++)
              T2(fqx, i, k) = R(0.5) * (WVEL(ru, i, k, j)) * (A3(w, i, k, j) + A3(w, i - 1, k, j));
          if (i == ids + 2)
            for (int k = kts + 1; k <= ktf + 1; k++) {
              wreal vel = WVEL(ru, i, k, j);
              T2(fqx, i, k) = vel * flux3(A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), vel, ts);
            }
        }
      if (degrade_xe)
        for (int i = i_end_f + 1; i <= i_end + 1; i++) {
          if (i == ide - 1)
            for (int k = kts + 1; k <= ktf + 1; k++)
              T2(fqx, i, k) = R(0.5) * (WVEL(ru, i, k, j)) * (A3(w, i, k, j) + A3(w, i - 1, k, j));
          if (i == ide - 2)
            for (int k = kts + 1; k <= ktf + 1; k++) {
              wreal vel = WVEL(ru, i, k, j);
              T2(fqx, i, k) = vel * flux3(A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), vel, ts);
            }
        }
      for (int k = kts + 1; k <= ktf + 1; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal mrdx = A2(msftx, i, j) * rdx;
          A3(tend, i, k, j) = A3(tend, i, k, j) - mrdx * (T2(fqx, i + 1, k) - T2(fqx, i, k));
        }
    }
  }
  {
    int i_end = IMIN(ite, ide - 1), j_end = IMIN(jte, jde - 1);
    /* open boundaries: the lid level uses the extrapolated face velocity and
     * difference, written out explicitly as module_advect_em.F:5703-5716 does */
    if (c->open_xs && its == ids)
      for (int j = jts; j <= j_end; j++)
        for (int k = kts + 1; k <= ktf + 1; k++) {
          wreal uw, dru;
          if (k <= ktf) {
            uw = R(0.5) * (A1(fzm, k) * (A3(ru, its, k, j) + A3(ru, its + 1, k, j)) + A1(fzp, k) * (A3(ru, its, k - 1, j) + A3(ru, its + 1, k - 1, j)));
            dru = A1(fzm, k) * (A3(ru, its + 1, k, j) - A3(ru, its, k, j)) + A1(fzp, k) * (A3(ru, its + 1, k - 1, j) - A3(ru, its, k - 1, j));
          } else {
            uw = R(0.5) * ((R(2.) - A1(fzm, k - 1)) * (A3(ru, its, k - 1, j) + A3(ru, its + 1, k - 1, j)) - A1(fzp, k - 1) * (A3(ru, its, k - 2, j) + A3(ru, its + 1, k - 2, j)));
            dru = (R(2.) - A1(fzm, k - 1)) * (A3(ru, its + 1, k - 1, j) - A3(ru, its, k - 1, j)) - A1(fzp, k - 1) * (A3(ru, its + 1, k - 2, j) - A3(ru, its, k - 2, j));
          }
          wreal ub = uw < R(0.) ? uw : R(0.);
          A3(tend, its, k, j) = A3(tend, its, k, j) - rdx * (ub * (A3(w_old, its + 1, k, j) - A3(w_old, its, k, j)) + A3(w, its, k, j) * dru);
        }
    if (c->open_xe && ite == ide)
      for (int j = jts; j <= j_end; j++)
        for (int k = kts + 1; k <= ktf + 1; k++) {
          wreal uw, dru;
          if (k <= ktf) {
            uw = R(0.5) * (A1(fzm, k) * (A3(ru, ite - 1, k, j) + A3(ru, ite, k, j)) + A1(fzp, k) * (A3(ru, ite - 1, k - 1, j) + A3(ru, ite, k - 1, j)));
            dru = A1(fzm, k) * (A3(ru, ite, k, j) - A3(ru, ite - 1, k, j)) + A1(fzp, k) * (A3(ru, ite, k - 1, j) - A3(ru, ite - 1, k - 1, j));
          } else {
            uw = R(0.5) * ((R(2.) - A1(fzm, k - 1)) * (A3(ru, ite - 1, k - 1, j) + A3(ru, ite, k - 1, j)) - A1(fzp, k - 1) * (A3(ru, ite - 1, k - 2, j) + A3(ru, ite, k - 2, j)));
            dru = (R(2.) - A1(fzm, k - 1)) * (A3(ru, ite, k - 1, j) - A3(ru, ite - 1, k - 1, j)) - A1(fzp, k - 1) * (A3(ru, ite, k - 2, j) - A3(ru, ite - 1, k - 2, j));
          }
          wreal ub = uw > R(0.) ? uw : R(0.);
          A3(tend, i_end, k, j) = A3(tend, i_end, k, j) - rdx * (ub * (A3(w_old, i_end, k, j) - A3(w_old, i_end - 1, k, j)) + A3(w, i_end, k, j) * dru);
        }
    if (c->open_ys && jts == jds)
      for (int i = its; i <= i_end; i++)
        for (int k = kts + 1; k <= ktf + 1; k++) {
          wreal vw, drv;
          if (k <= ktf) {
            vw = R(0.5) * (A1(fzm, k) * (A3(rv, i, k, jts) + A3(rv, i, k, jts + 1)) + A1(fzp, k) * (A3(rv, i, k - 1, jts) + A3(rv, i, k - 1, jts + 1)));
            drv = A1(fzm, k) * (A3(rv, i, k, jts + 1) - A3(rv, i, k, jts)) + A1(fzp, k) * (A3(rv, i, k - 1, jts + 1) - A3(rv, i, k - 1, jts));
          } else {
            vw = R(0.5) * ((R(2.) - A1(fzm, k - 1)) * (A3(rv, i, k - 1, jts) + A3(rv, i, k - 1, jts + 1)) - A1(fzp, k - 1) * (A3(rv, i, k - 2, jts) + A3(rv, i, k - 2, jts + 1)));
            drv = (R(2.) - A1(fzm, k - 1)) * (A3(rv, i, k - 1, jts + 1) - A3(rv, i, k - 1, jts)) - A1(fzp, k - 1) * (A3(rv, i, k - 2, jts + 1) - A3(rv, i, k - 2, jts));
          }
          wreal vb = vw < R(0.) ? vw : R(0.);
          A3(tend, i, k, jts) = A3(tend, i, k, jts) - rdy * (vb * (A3(w_old, i, k, jts + 1) - A3(w_old, i, k, jts)) + A3(w, i, k, jts) * drv);
        }
    if (c->open_ye && jte == jde)
      for (int i = its; i <= i_end; i++)
        for (int k = kts + 1; k <= ktf + 1; k++) {
          wreal vw, drv;
          if (k <= ktf) {
            vw = R(0.5) * (A1(fzm, k) * (A3(rv, i, k, jte - 1) + A3(rv, i, k, jte)) + A1(fzp, k) * (A3(rv, i, k - 1, jte - 1) + A3(rv, i, k - 1, jte)));
            drv = A1(fzm, k) * (A3(rv, i, k, jte) - A3(rv, i, k, jte - 1)) + A1(fzp, k) * (A3(rv, i, k - 1, jte) - A3(rv, i, k - 1, jte - 1));
          } else {
            vw = R(0.5) * ((R(2.) - A1(fzm, k - 1)) * (A3(rv, i, k - 1, jte - 1) + A3(rv, i, k - 1, jte)) - A1(fzp, k - 1) * (A3(rv, i, k - 2, jte - 1) + A3(rv, i, k - 2, jte)));
            drv = (R(2.) - A1(fzm, k - 1)) * (A3(rv, i, k - 1, jte) - A3(rv, i, k - 1, jte - 1)) - A1(fzp, k - 1) * (A3(rv, i, k - 2, jte) - A3(rv, i, k - 2, jte - 1));
          }
          wreal vb = vw > R(0.) ? vw : R(0.);
          A3(tend, i, k, j_end) = A3(tend, i, k, j_end) - rdy * (vb * (A3(w_old, i, k, j_end) - A3(w_old, i, k, j_end - 1)) + A3(w, i, k, j_end) * drv);
        }
  }
  {
    int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
    for (int i = i_start; i <= i_end; i++) { T2(vflux, i, kts) = R(0.); T2(vflux, i, kte) = R(0.); }
    for (int j = j_start; j <= j_end; j++) {
      for (int k = kts + 2; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal vel = R(0.5) * (A3(rom, i, k, j) + A3(rom, i, k - 1, j));
          T2(vflux, i, k) = vel * flux3(A3(w, i, k - 2, j), A3(w, i, k - 1, j), A3(w, i, k, j), A3(w, i, k + 1, j), -vel, ts);
        }
      for (int i = i_start; i <= i_end; i++) {
        int k = kts + 1;
        T2(vflux, i, k) = R(0.25) * (A3(rom, i, k, j) + A3(rom, i, k - 1, j)) * (A3(w, i, k, j) + A3(w, i, k - 1, j));
        k = ktf + 1;
        T2(vflux, i, k) = R(0.25) * (A3(rom, i, k, j) + A3(rom, i, k - 1, j)) * (A3(w, i, k, j) + A3(w, i, k - 1, j));
      }
      for (int k = kts + 1; k <= ktf; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(tend, i, k, j) = A3(tend, i, k, j) - A1(rdzu, k) * (T2(vflux, i, k + 1) - T2(vflux, i, k));
      {
        int k = ktf + 1;
        for (int i = i_start; i <= i_end; i++)
          A3(tend, i, k, j) = A3(tend, i, k, j) + R(2.) * A1(rdzu, k - 1) * (T2(vflux, i, k));
      }
    }
  }
#undef WVEL
  free(fqx.p); free(fqy0.p); free(fqy1.p); free(vflux.p);
  return 0;
}
