/*
 * Oracle: acoustic small step.  TEST INFRASTRUCTURE.
 *   small_step_prep   dyn_em/module_small_step_em.F:16-290
 *   small_step_finish dyn_em/module_small_step_em.F:295-434
 *   calc_p_rho        dyn_em/module_small_step_em.F:438-568
 *   calc_coef_w       dyn_em/module_small_step_em.F:570-652
 *   advance_uv        dyn_em/module_small_step_em.F:654-967
 *   advance_mu_t      dyn_em/module_small_step_em.F:969-1175
 *   advance_w         dyn_em/module_small_step_em.F:1178-1469
 *   sumflux           dyn_em/module_small_step_em.F:1473-1633
 *   calc_mu_uv_1      dyn_em/module_big_step_utilities_em.F:184-321
 * Arguments as bound at the solve_em.F call sites (1089-1757): ww_save = ww1,
 * t_1 (in calc_p_rho/advance_mu_t/advance_w) = t_save, mu = mu_2, mut = muts
 * (calc_p_rho), w_save/ph_save/mu_save/t_2save/c2a/pm1/a/alpha/gamma = solve_em locals.
 */
#include <stdlib.h>
#include "oracle.h"

void o_small_step_prep(ostate *s, const dims_t *d, int rk_step) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  int k_start = kts, k_end = IMIN(kte, kde - 1), i_endu = ite, j_endv = jte;
  wreal *c1h = s->c1h, *c2h = s->c2h, *c1f = s->c1f, *c2f = s->c2f;
  if (rk_step == 1) {
    for (int j = j_start; j <= j_end; j++)
      for (int i = i_start; i <= i_end; i++) {
        A2(s->mu_1, i, j) = A2(s->mu_2, i, j);
        A3(s->ww1, i, kde, j) = R(0.);
        A3(s->ww1, i, 1, j) = R(0.);
        A2(s->mudf, i, j) = R(0.);
      }
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_endu; i++) A3(s->u_1, i, k, j) = A3(s->u_2, i, k, j);
    for (int j = j_start; j <= j_endv; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) A3(s->v_1, i, k, j) = A3(s->v_2, i, k, j);
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) A3(s->t_1, i, k, j) = A3(s->t_2, i, k, j);
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= IMIN(kde, kte); k++)
        for (int i = i_start; i <= i_end; i++) {
          A3(s->w_1, i, k, j) = A3(s->w_2, i, k, j);
          A3(s->ph_1, i, k, j) = A3(s->ph_2, i, k, j);
        }
    for (int j = j_start; j <= j_end; j++) {
      for (int i = i_start; i <= i_end; i++) A2(s->muts, i, j) = A2(s->mub, i, j) + A2(s->mu_2, i, j);
      for (int i = i_start; i <= i_endu; i++) A2(s->muus, i, j) = A2(s->muu, i, j);
    }
    for (int j = j_start; j <= j_endv; j++)
      for (int i = i_start; i <= i_end; i++) A2(s->muvs, i, j) = A2(s->muv, i, j);
    for (int j = j_start; j <= j_end; j++)
      for (int i = i_start; i <= i_end; i++) {
        A2(s->mu_save, i, j) = A2(s->mu_2, i, j);
        A2(s->mu_2, i, j) = R(0.);
      }
  } else {
    for (int j = j_start; j <= j_end; j++) {
      for (int i = i_start; i <= i_end; i++) A2(s->muts, i, j) = A2(s->mub, i, j) + A2(s->mu_1, i, j);
      for (int i = i_start; i <= i_endu; i++)
        A2(s->muus, i, j) = R(0.5) * (((A2(s->mub, i, j) + A2(s->mu_1, i, j)) + A2(s->mub, i - 1, j)) + A2(s->mu_1, i - 1, j));
    }
    for (int j = j_start; j <= j_endv; j++)
      for (int i = i_start; i <= i_end; i++)
        A2(s->muvs, i, j) = R(0.5) * (((A2(s->mub, i, j) + A2(s->mu_1, i, j)) + A2(s->mub, i, j - 1)) + A2(s->mu_1, i, j - 1));
    for (int j = j_start; j <= j_end; j++)
      for (int i = i_start; i <= i_end; i++) {
        A2(s->mu_save, i, j) = A2(s->mu_2, i, j);
        A2(s->mu_2, i, j) = A2(s->mu_1, i, j) - A2(s->mu_2, i, j);
      }
  }
  for (int j = j_start; j <= j_end; j++)
    for (int i = i_start; i <= i_end; i++) {
      A3(s->ww1, i, kde, j) = R(0.);
      A3(s->ww1, i, 1, j) = R(0.);
    }
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++)
        A3(s->c2a, i, k, j) = K_CPOVCV * (A3(s->pb, i, k, j) + A3(s->p, i, k, j)) / A3(s->alt, i, k, j);
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_endu; i++) {
        A3(s->u_save, i, k, j) = A3(s->u_2, i, k, j);
        A3(s->u_2, i, k, j) = ((A1(c1h, k) * A2(s->muus, i, j) + A1(c2h, k)) * A3(s->u_1, i, k, j) -
                               (A1(c1h, k) * A2(s->muu, i, j) + A1(c2h, k)) * A3(s->u_2, i, k, j)) / A2(s->msfuy, i, j);
      }
  for (int j = j_start; j <= j_endv; j++)
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++) {
        A3(s->v_save, i, k, j) = A3(s->v_2, i, k, j);
        A3(s->v_2, i, k, j) = ((A1(c1h, k) * A2(s->muvs, i, j) + A1(c2h, k)) * A3(s->v_1, i, k, j) -
                               (A1(c1h, k) * A2(s->muv, i, j) + A1(c2h, k)) * A3(s->v_2, i, k, j)) * A2(s->msfvx_inv, i, j);
      }
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++) {
        A3(s->t_save, i, k, j) = A3(s->t_2, i, k, j);
        A3(s->t_2, i, k, j) = (A1(c1h, k) * A2(s->muts, i, j) + A1(c2h, k)) * A3(s->t_1, i, k, j) -
                              (A1(c1h, k) * A2(s->mut, i, j) + A1(c2h, k)) * A3(s->t_2, i, k, j);
      }
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= kde; k++)
      for (int i = i_start; i <= i_end; i++) {
        A3(s->w_save, i, k, j) = A3(s->w_2, i, k, j);
        A3(s->w_2, i, k, j) = ((A1(c1f, k) * A2(s->muts, i, j) + A1(c2f, k)) * A3(s->w_1, i, k, j) -
                               (A1(c1f, k) * A2(s->mut, i, j) + A1(c2f, k)) * A3(s->w_2, i, k, j)) / A2(s->msfty, i, j);
        A3(s->ph_save, i, k, j) = A3(s->ph_2, i, k, j);
        A3(s->ph_2, i, k, j) = A3(s->ph_1, i, k, j) - A3(s->ph_2, i, k, j);
      }
  for (int j = j_start; j <= j_end; j++)
    for (int k = k_start; k <= kde; k++)
      for (int i = i_start; i <= i_end; i++) A3(s->ww1, i, k, j) = A3(s->ww, i, k, j);
}

void o_small_step_finish(ostate *s, const dims_t *d, int nsmall, wreal dts, int rk_step, int rk_order) {
  const int ide = d->ide, jde = d->jde, kds = d->kds, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  int i_endu = ite, j_endv = jte;
  wreal *c1h = s->c1h, *c2h = s->c2h, *c1f = s->c1f, *c2f = s->c2f;
  for (int j = j_start; j <= j_endv; j++)
    for (int k = kds; k <= kde - 1; k++)
      for (int i = i_start; i <= i_end; i++)
        A3(s->v_2, i, k, j) = (A2(s->msfvx, i, j) * A3(s->v_2, i, k, j) + A3(s->v_save, i, k, j) * (A1(c1h, k) * A2(s->muv, i, j) + A1(c2h, k))) /
                              (A1(c1h, k) * A2(s->muvs, i, j) + A1(c2h, k));
  for (int j = j_start; j <= j_end; j++)
    for (int k = kds; k <= kde - 1; k++)
      for (int i = i_start; i <= i_endu; i++)
        A3(s->u_2, i, k, j) = (A2(s->msfuy, i, j) * A3(s->u_2, i, k, j) + A3(s->u_save, i, k, j) * (A1(c1h, k) * A2(s->muu, i, j) + A1(c2h, k))) /
                              (A1(c1h, k) * A2(s->muus, i, j) + A1(c2h, k));
  for (int j = j_start; j <= j_end; j++)
    for (int k = kds; k <= kde; k++)
      for (int i = i_start; i <= i_end; i++) {
        A3(s->w_2, i, k, j) = (A2(s->msfty, i, j) * A3(s->w_2, i, k, j) + A3(s->w_save, i, k, j) * (A1(c1f, k) * A2(s->mut, i, j) + A1(c2f, k))) /
                              (A1(c1f, k) * A2(s->muts, i, j) + A1(c2f, k));
        A3(s->ph_2, i, k, j) = A3(s->ph_2, i, k, j) + A3(s->ph_save, i, k, j);
        A3(s->ww, i, k, j) = A3(s->ww, i, k, j) + A3(s->ww1, i, k, j);
      }
  if (rk_step < rk_order) {
    for (int j = j_start; j <= j_end; j++)
      for (int k = kds; k <= kde - 1; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(s->t_2, i, k, j) = (A3(s->t_2, i, k, j) + A3(s->t_save, i, k, j) * (A1(c1h, k) * A2(s->mut, i, j) + A1(c2h, k))) /
                                (A1(c1h, k) * A2(s->muts, i, j) + A1(c2h, k));
  } else {
    for (int j = j_start; j <= j_end; j++)
      for (int k = kds; k <= kde - 1; k++)
        for (int i = i_start; i <= i_end; i++)
          A3(s->t_2, i, k, j) = (A3(s->t_2, i, k, j) - dts * (wreal)nsmall * (A1(c1h, k) * A2(s->mut, i, j) + A1(c2h, k)) * A3(s->h_diabatic, i, k, j) +
                                 A3(s->t_save, i, k, j) * (A1(c1h, k) * A2(s->mut, i, j) + A1(c2h, k))) /
                                (A1(c1h, k) * A2(s->muts, i, j) + A1(c2h, k));
  }
  for (int j = j_start; j <= j_end; j++)
    for (int i = i_start; i <= i_end; i++) A2(s->mu_2, i, j) = A2(s->mu_2, i, j) + A2(s->mu_save, i, j);
}

void o_calc_p_rho(ostate *s, const cfg_t *c, const dims_t *d, int step) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  int k_start = kts, k_end = IMIN(kte, kde - 1);
  wreal *al = s->al, *p = s->p, *ph = s->ph_2, *alt = s->alt, *t_2 = s->t_2, *t_1 = s->t_save, *c2a = s->c2a;
  wreal *mu = s->mu_2, *mut = s->muts, *c1h = s->c1h, *c2h = s->c2h;
  if (c->non_hydrostatic) {
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) {
          A3(al, i, k, j) = -R(1.) / (A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) *
                            (A3(alt, i, k, j) * (A1(c1h, k) * A2(mu, i, j)) + A1(s->rdnw, k) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)));
          A3(p, i, k, j) = A3(c2a, i, k, j) * (A3(alt, i, k, j) * (A3(t_2, i, k, j) - (A1(c1h, k) * A2(mu, i, j)) * A3(t_1, i, k, j)) /
                                                   ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (K_T0 + A3(t_1, i, k, j))) -
                                               A3(al, i, k, j));
        }
  } else {
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) {
          A3(p, i, k, j) = A2(mu, i, j) * A1(s->c3h, k);
          A3(al, i, k, j) = A3(alt, i, k, j) * (A3(t_2, i, k, j) - (A1(c1h, k) * A2(mu, i, j)) * A3(t_1, i, k, j)) /
                                ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (K_T0 + A3(t_1, i, k, j))) -
                            A3(p, i, k, j) / A3(c2a, i, k, j);
          A3(ph, i, k + 1, j) = A3(ph, i, k, j) - A1(s->dnw, k) * ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * A3(al, i, k, j) +
                                                                   (A1(c1h, k) * A2(mu, i, j)) * A3(alt, i, k, j));
        }
  }
  if (step == 0) {
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) A3(s->pm1, i, k, j) = A3(p, i, k, j);
  } else {
    wreal smdiv = (wreal)c->smdiv;
    for (int j = j_start; j <= j_end; j++)
      for (int k = k_start; k <= k_end; k++)
        for (int i = i_start; i <= i_end; i++) {
          wreal ptmp = A3(p, i, k, j);
          A3(p, i, k, j) = A3(p, i, k, j) + smdiv * (A3(p, i, k, j) - A3(s->pm1, i, k, j));
          A3(s->pm1, i, k, j) = ptmp;
        }
  }
}

/* a, alpha, gamma are WREAL arrays in the oracle (calc_coef_w is REAL in the reference) */
void o_calc_coef_w(ostate *s, const cfg_t *c, const dims_t *d, wreal dts) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  wreal *a = (wreal *)s->a, *alpha = (wreal *)s->alpha, *gamma = (wreal *)s->gamma;
  wreal *c2a = s->c2a, *cqw = s->cqw, *mut = s->mut;
  wreal *c1h = s->c1h, *c2h = s->c2h, *c1f = s->c1f, *c2f = s->c2f, *rdn = s->rdn, *rdnw = s->rdnw;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  wreal lid_flag = c->top_lid ? R(0.) : R(1.);
  wreal epssm = (wreal)c->epssm;
  for (int j = j_start; j <= j_end; j++) {
    int k = kde - 1;
    for (int i = i_start; i <= i_end; i++) {
      wreal cof = (R(.5) * dts * K_G * (R(1.) + epssm));
      cof = cof * cof;
      A3(a, i, 2, j) = R(0.);
      A3(a, i, kde, j) = -R(2.) * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * A3(c2a, i, kde - 1, j) * lid_flag /
                         ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)));
      A3(gamma, i, 1, j) = R(0.);
    }
    for (int kk = 3; kk <= kde - 1; kk++) {
      k = kk - 1;
      for (int i = i_start; i <= i_end; i++) {
        wreal cof = (R(.5) * dts * K_G * (R(1.) + epssm));
        cof = cof * cof;
        A3(a, i, kk, j) = -A3(cqw, i, kk, j) * cof * A1(rdn, kk) * A1(rdnw, kk - 1) * A3(c2a, i, kk - 1, j) /
                          ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)));
      }
    }
    for (k = 2; k <= kde - 1; k++)
      for (int i = i_start; i <= i_end; i++) {
        wreal cof = (R(.5) * dts * K_G * (R(1.) + epssm));
        cof = cof * cof;
        wreal b = R(1.) + A3(cqw, i, k, j) * cof * A1(rdn, k) *
                              (A1(rdnw, k) * A3(c2a, i, k, j) / ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k))) +
                               A1(rdnw, k - 1) * A3(c2a, i, k - 1, j) / ((A1(c1h, k - 1) * A2(mut, i, j) + A1(c2h, k - 1)) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k))));
        wreal cc = -A3(cqw, i, k, j) * cof * A1(rdn, k) * A1(rdnw, k) * A3(c2a, i, k, j) /
                   ((A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * (A1(c1f, k + 1) * A2(mut, i, j) + A1(c2f, k + 1)));
        A3(alpha, i, k, j) = R(1.) / (b - A3(a, i, k, j) * A3(gamma, i, k - 1, j));
        A3(gamma, i, k, j) = cc * A3(alpha, i, k, j);
      }
    k = kde;
    for (int i = i_start; i <= i_end; i++) {
      wreal cof = (R(.5) * dts * K_G * (R(1.) + epssm));
      cof = cof * cof;
      wreal b = R(1.) + R(2.) * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * A3(c2a, i, kde - 1, j) /
                            ((A1(c1h, k - 1) * A2(mut, i, j) + A1(c2h, k - 1)) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)));
      wreal cc = R(0.);
      A3(alpha, i, kde, j) = R(1.) / (b - A3(a, i, kde, j) * A3(gamma, i, kde - 1, j));
      A3(gamma, i, kde, j) = cc * A3(alpha, i, kde, j);
    }
  }
}

void o_advance_uv(ostate *s, const cfg_t *c, const dims_t *d, wreal dts) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  int i_start, i_end, j_start, j_end, k_start, k_end, i_endu, j_endv;
  if (c->nested || c->specified) {
    /* module_small_step_em.F:734-747: the spec zone is left to spec_bdyupdate */
    const int sz = c->spec_zone;
    i_start = IMAX(its, ids + sz); i_end = IMIN(ite, ide - sz - 1); j_start = IMAX(jts, jds + sz);
    j_end = IMIN(jte, jde - sz - 1); k_start = kts; k_end = IMIN(kte, kde - 1); i_endu = IMIN(ite, ide - sz);
    j_endv = IMIN(jte, jde - sz);
    if (c->periodic_x) { i_start = its; i_end = IMIN(ite, ide - 1); i_endu = ite; }
  } else {
    i_start = its; i_end = IMIN(ite, ide - 1); j_start = jts; j_end = IMIN(jte, jde - 1);
    k_start = kts; k_end = kte - 1; i_endu = ite; j_endv = jte;
  }
  int i_start_up = i_start, i_end_up = i_endu, j_start_vp = j_start, j_end_vp = j_endv;
  if ((c->open_xs || c->symmetric_xs) && its == ids) i_start_up = i_start_up + 1;
  if ((c->open_xe || c->symmetric_xe) && ite == ide) i_end_up = i_end_up - 1;
  if ((c->open_ys || c->symmetric_ys || c->polar) && jts == jds) j_start_vp = j_start_vp + 1;
  if ((c->open_ye || c->symmetric_ye || c->polar) && jte == jde) j_end_vp = j_end_vp - 1;
  int i_start_u_tend = i_start, i_end_u_tend = i_endu, j_start_v_tend = j_start, j_end_v_tend = j_endv;
  if (c->symmetric_xs && its == ids) i_start_u_tend++;
  if (c->symmetric_xe && ite == ide) i_end_u_tend--;
  if (c->symmetric_ys && jts == jds) j_start_v_tend++;
  if (c->symmetric_ye && jte == jde) j_end_v_tend--;
  wreal rdx = s->rdx, rdy = s->rdy, dx = R(1.) / rdx, dy = R(1.) / rdy;
  wreal cf1 = s->cf1, cf2 = s->cf2, cf3 = s->cf3, emdiv = (wreal)c->emdiv;
  wreal *u = s->u_2, *v = s->v_2, *p = s->p, *pb = s->pb, *ph = s->ph_2, *php = s->php, *alt = s->alt, *al = s->al;
  wreal *mu = s->mu_2, *mudf = s->mudf, *c1h = s->c1h, *c2h = s->c2h;
  int nk = NK_;
  wreal *dpn = (wreal *)calloc((size_t)nk + 1, sizeof(wreal)), *dpxy = (wreal *)calloc((size_t)nk + 1, sizeof(wreal));
#define DPN(k) dpn[(k) - d->kms]
#define DPXY(k) dpxy[(k) - d->kms]
  for (int j = j_start; j <= j_end; j++) {
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start_u_tend; i <= i_end_u_tend; i++) A3(u, i, k, j) = A3(u, i, k, j) + dts * A3(s->ru_tend, i, k, j);
    for (int i = i_start_up; i <= i_end_up; i++) {
      wreal mudf_xy = -emdiv * dx * (A2(mudf, i, j) - A2(mudf, i - 1, j)) / A2(s->msfuy, i, j);
      for (int k = k_start; k <= k_end; k++)
        DPXY(k) = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * R(.5) * rdx * (A1(c1h, k) * A2(s->muu, i, j) + A1(c2h, k)) *
                  (((A3(ph, i, k + 1, j) - A3(ph, i - 1, k + 1, j)) + (A3(ph, i, k, j) - A3(ph, i - 1, k, j))) +
                   (A3(alt, i, k, j) + A3(alt, i - 1, k, j)) * (A3(p, i, k, j) - A3(p, i - 1, k, j)) +
                   (A3(al, i, k, j) + A3(al, i - 1, k, j)) * (A3(pb, i, k, j) - A3(pb, i - 1, k, j)));
      if (c->non_hydrostatic) {
        DPN(1) = R(.5) * (cf1 * (A3(p, i, 1, j) + A3(p, i - 1, 1, j)) + cf2 * (A3(p, i, 2, j) + A3(p, i - 1, 2, j)) +
                          cf3 * (A3(p, i, 3, j) + A3(p, i - 1, 3, j)));
        DPN(kde) = R(0.);
        if (c->top_lid)
          DPN(kde) = R(.5) * (cf1 * (A3(p, i - 1, kde - 1, j) + A3(p, i, kde - 1, j)) + cf2 * (A3(p, i - 1, kde - 2, j) + A3(p, i, kde - 2, j)) +
                              cf3 * (A3(p, i - 1, kde - 3, j) + A3(p, i, kde - 3, j)));
        for (int k = k_start + 1; k <= k_end; k++)
          DPN(k) = R(.5) * (A1(s->fnm, k) * (A3(p, i, k, j) + A3(p, i - 1, k, j)) + A1(s->fnp, k) * (A3(p, i, k - 1, j) + A3(p, i - 1, k - 1, j)));
        for (int k = k_start; k <= k_end; k++)
          DPXY(k) = DPXY(k) + (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * rdx * (A3(php, i, k, j) - A3(php, i - 1, k, j)) *
                                  (A1(s->rdnw, k) * (DPN(k + 1) - DPN(k)) - R(.5) * ((A1(c1h, k) * A2(mu, i - 1, j)) + (A1(c1h, k) * A2(mu, i, j))));
      }
      for (int k = k_start; k <= k_end; k++)
        A3(u, i, k, j) = A3(u, i, k, j) - dts * A3(s->cqu, i, k, j) * DPXY(k) + (A1(c1h, k) * mudf_xy);
    }
  }
  for (int j = j_start_v_tend; j <= j_end_v_tend; j++) {
    for (int k = k_start; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++) A3(v, i, k, j) = A3(v, i, k, j) + dts * A3(s->rv_tend, i, k, j);
    if (j >= j_start_vp && j <= j_end_vp)
      for (int i = i_start; i <= i_end; i++) {
        wreal mudf_xy = -emdiv * dy * (A2(mudf, i, j) - A2(mudf, i, j - 1)) * A2(s->msfvx_inv, i, j);
        for (int k = k_start; k <= k_end; k++)
          DPXY(k) = (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * R(.5) * rdy * (A1(c1h, k) * A2(s->muv, i, j) + A1(c2h, k)) *
                    (((A3(ph, i, k + 1, j) - A3(ph, i, k + 1, j - 1)) + (A3(ph, i, k, j) - A3(ph, i, k, j - 1))) +
                     (A3(alt, i, k, j) + A3(alt, i, k, j - 1)) * (A3(p, i, k, j) - A3(p, i, k, j - 1)) +
                     (A3(al, i, k, j) + A3(al, i, k, j - 1)) * (A3(pb, i, k, j) - A3(pb, i, k, j - 1)));
        if (c->non_hydrostatic) {
          DPN(1) = R(.5) * (cf1 * (A3(p, i, 1, j) + A3(p, i, 1, j - 1)) + cf2 * (A3(p, i, 2, j) + A3(p, i, 2, j - 1)) +
                            cf3 * (A3(p, i, 3, j) + A3(p, i, 3, j - 1)));
          DPN(kde) = R(0.);
          if (c->top_lid)
            DPN(kde) = R(.5) * (cf1 * (A3(p, i, kde - 1, j - 1) + A3(p, i, kde - 1, j)) + cf2 * (A3(p, i, kde - 2, j - 1) + A3(p, i, kde - 2, j)) +
                                cf3 * (A3(p, i, kde - 3, j - 1) + A3(p, i, kde - 3, j)));
          for (int k = k_start + 1; k <= k_end; k++)
            DPN(k) = R(.5) * (A1(s->fnm, k) * (A3(p, i, k, j) + A3(p, i, k, j - 1)) + A1(s->fnp, k) * (A3(p, i, k - 1, j) + A3(p, i, k - 1, j - 1)));
          for (int k = k_start; k <= k_end; k++)
            DPXY(k) = DPXY(k) + (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * rdy * (A3(php, i, k, j) - A3(php, i, k, j - 1)) *
                                    (A1(s->rdnw, k) * (DPN(k + 1) - DPN(k)) - R(.5) * ((A1(c1h, k) * A2(mu, i, j - 1)) + (A1(c1h, k) * A2(mu, i, j))));
        }
        for (int k = k_start; k <= k_end; k++)
          A3(v, i, k, j) = A3(v, i, k, j) - dts * A3(s->cqv, i, k, j) * DPXY(k) + (A1(c1h, k) * mudf_xy);
      }
  }
#undef DPN
#undef DPXY
  free(dpn); free(dpxy);
}

void o_advance_mu_t(ostate *s, const cfg_t *c, const dims_t *d, wreal dts) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  (void)kde;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  int k_start = kts, k_end = kte - 1;
  if (!c->periodic_x && (c->specified || c->nested)) { i_start = IMAX(its, ids + 1); i_end = IMIN(ite, ide - 2); }
  if (c->specified || c->nested) { j_start = IMAX(jts, jds + 1); j_end = IMIN(jte, jde - 2); }
  wreal *ww = s->ww, *ww_1 = s->ww1, *u = s->u_2, *u_1 = s->u_save, *v = s->v_2, *v_1 = s->v_save;
  wreal *mu = s->mu_2, *mut = s->mut, *muave = s->muave, *muts = s->muts, *muu = s->muu, *muv = s->muv, *mudf = s->mudf;
  wreal *t = s->t_2, *t_1 = s->t_save, *t_ave = s->t_2save, *ft = s->t_tend, *mu_tend = s->mu_tend;
  wreal *c1h = s->c1h, *c2h = s->c2h, rdx = s->rdx, rdy = s->rdy, epssm = (wreal)c->epssm;
  int nk = NK_;
  wreal *dvdxi = (wreal *)calloc((size_t)nk + 1, sizeof(wreal)), *wdtn = (wreal *)calloc((size_t)nk + 1, sizeof(wreal));
#define DV(k) dvdxi[(k) - d->kms]
#define WDTN(k) wdtn[(k) - d->kms]
  for (int j = j_start; j <= j_end; j++)
    for (int i = i_start; i <= i_end; i++) {
      wreal dmdt = R(0.);
      for (int k = k_start; k <= k_end; k++) {
        DV(k) = A2(s->msftx, i, j) * A2(s->msfty, i, j) *
                (rdy * ((A3(v, i, k, j + 1) + (A1(c1h, k) * A2(muv, i, j + 1) + A1(c2h, k)) * A3(v_1, i, k, j + 1) * A2(s->msfvx_inv, i, j + 1)) -
                        (A3(v, i, k, j) + (A1(c1h, k) * A2(muv, i, j) + A1(c2h, k)) * A3(v_1, i, k, j) * A2(s->msfvx_inv, i, j))) +
                 rdx * ((A3(u, i + 1, k, j) + (A1(c1h, k) * A2(muu, i + 1, j) + A1(c2h, k)) * A3(u_1, i + 1, k, j) / A2(s->msfuy, i + 1, j)) -
                        (A3(u, i, k, j) + (A1(c1h, k) * A2(muu, i, j) + A1(c2h, k)) * A3(u_1, i, k, j) / A2(s->msfuy, i, j))));
        dmdt = dmdt + A1(s->dnw, k) * DV(k);
      }
      A2(muave, i, j) = A2(mu, i, j);
      A2(mu, i, j) = A2(mu, i, j) + dts * (dmdt + A2(mu_tend, i, j));
      A2(mudf, i, j) = (dmdt + A2(mu_tend, i, j));
      A2(muts, i, j) = A2(mut, i, j) + A2(mu, i, j);
      A2(muave, i, j) = R(.5) * ((R(1.) + epssm) * A2(mu, i, j) + (R(1.) - epssm) * A2(muave, i, j));
      for (int kk = 2; kk <= k_end; kk++) {
        int k = kk - 1;
        A3(ww, i, kk, j) = A3(ww, i, kk - 1, j) - A1(s->dnw, kk - 1) * ((A1(c1h, k) * dmdt) + DV(kk - 1) + (A1(c1h, k) * A2(mu_tend, i, j))) / A2(s->msfty, i, j);
      }
      for (int k = 1; k <= k_end; k++) A3(ww, i, k, j) = A3(ww, i, k, j) - A3(ww_1, i, k, j);
    }
  for (int j = j_start; j <= j_end; j++)
    for (int k = 1; k <= k_end; k++)
      for (int i = i_start; i <= i_end; i++) {
        A3(t_ave, i, k, j) = A3(t, i, k, j);
        A3(t, i, k, j) = A3(t, i, k, j) + A2(s->msfty, i, j) * dts * A3(ft, i, k, j);
      }
  for (int j = j_start; j <= j_end; j++)
    for (int i = i_start; i <= i_end; i++) {
      WDTN(1) = R(0.);
      WDTN(d->kde) = R(0.);
      for (int k = 2; k <= k_end; k++)
        WDTN(k) = A3(ww, i, k, j) * (A1(s->fnm, k) * A3(t_1, i, k, j) + A1(s->fnp, k) * A3(t_1, i, k - 1, j));
      for (int k = 1; k <= k_end; k++)
        A3(t, i, k, j) = A3(t, i, k, j) -
            dts * A2(s->msfty, i, j) *
                (A2(s->msftx, i, j) *
                     (R(.5) * rdy * (A3(v, i, k, j + 1) * (A3(t_1, i, k, j + 1) + A3(t_1, i, k, j)) - A3(v, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i, k, j - 1))) +
                      R(.5) * rdx * (A3(u, i + 1, k, j) * (A3(t_1, i + 1, k, j) + A3(t_1, i, k, j)) - A3(u, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i - 1, k, j)))) +
                 A1(s->rdnw, k) * (WDTN(k + 1) - WDTN(k)));
    }
#undef DV
#undef WDTN
  free(dvdxi); free(wdtn);
}

void o_advance_w(ostate *s, const cfg_t *c, const dims_t *d, wreal dts) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  (void)kts;
  int i_start = its, i_end = IMIN(ite, ide - 1), j_start = jts, j_end = IMIN(jte, jde - 1);
  int k_end = kte - 1;
  if (!c->periodic_x && (c->specified || c->nested)) { i_start = IMAX(its, ids + 1); i_end = IMIN(ite, ide - 2); }
  if (c->specified || c->nested) { j_start = IMAX(jts, jds + 1); j_end = IMIN(jte, jde - 2); }
  wreal *w = s->w_2, *rw_tend = s->rw_tend, *ww = s->ww, *w_save = s->w_save, *u = s->u_2, *v = s->v_2;
  wreal *mut = s->mut, *muave = s->muave, *muts = s->muts, *t_2ave = s->t_2save, *t_2 = s->t_2, *t_1 = s->t_save;
  wreal *ph = s->ph_2, *ph_1 = s->ph_save, *phb = s->phb, *ph_tend = s->ph_tend, *ht = s->ht, *c2a = s->c2a, *cqw = s->cqw, *alt = s->alt;
  wreal *a = (wreal *)s->a, *alpha = (wreal *)s->alpha, *gamma = (wreal *)s->gamma;
  wreal *c1h = s->c1h, *c2h = s->c2h, *c1f = s->c1f, *c2f = s->c2f, *rdn = s->rdn, *rdnw = s->rdnw;
  wreal rdx = s->rdx, rdy = s->rdy, cf1 = s->cf1, cf2 = s->cf2, cf3 = s->cf3, epssm = (wreal)c->epssm;
  wreal dampmag = dts * (wreal)c->dampcoef, hdepth = (wreal)c->zdamp;
  wreal pi = R(4.) * (wreal)atan((double)R(1.));
  int nk = NK_;
  wreal *rhs = (wreal *)calloc((size_t)nk + 1, sizeof(wreal)), *wdwn = (wreal *)calloc((size_t)nk + 1, sizeof(wreal));
#define RHS(k) rhs[(k) - d->kms]
#define WDWN(k) wdwn[(k) - d->kms]
  for (int j = j_start; j <= j_end; j++)
    for (int i = i_start; i <= i_end; i++) {
      RHS(1) = R(0.);
      wreal msft_inv = R(1.) / A2(s->msfty, i, j);
      for (int k = 1; k <= k_end; k++) {
        A3(t_2ave, i, k, j) = R(.5) * ((R(1.) + epssm) * A3(t_2, i, k, j) + (R(1.) - epssm) * A3(t_2ave, i, k, j));
        A3(t_2ave, i, k, j) = (A3(t_2ave, i, k, j) + (A1(c1h, k) * A2(muave, i, j)) * K_T0) /
                              ((A1(c1h, k) * A2(muts, i, j) + A1(c2h, k)) * (K_T0 + A3(t_1, i, k, j)));
        RHS(k + 1) = dts * (A3(ph_tend, i, k + 1, j) + R(.5) * K_G * (R(1.) - epssm) * A3(w, i, k + 1, j));
      }
      if (c->phi_adv_z == 2) {
        for (int k = 1; k <= k_end; k++)
          WDWN(k + 1) = A1(rdnw, k) * (A3(ph_1, i, k + 1, j) - A3(ph_1, i, k, j) + A3(phb, i, k + 1, j) - A3(phb, i, k, j));
        for (int k = 2; k <= k_end; k++)
          RHS(k) = RHS(k) - dts * A3(ww, i, k, j) * (A1(s->fnm, k) * WDWN(k + 1) + A1(s->fnp, k) * WDWN(k));
      } else {
        for (int k = 1; k <= k_end; k++)
          WDWN(k + 1) = R(.5) * (A3(ww, i, k + 1, j) + A3(ww, i, k, j)) * A1(rdnw, k) *
                        (A3(ph_1, i, k + 1, j) - A3(ph_1, i, k, j) + A3(phb, i, k + 1, j) - A3(phb, i, k, j));
        for (int k = 2; k <= k_end; k++)
          RHS(k) = RHS(k) - dts * (A1(s->fnm, k) * WDWN(k + 1) + A1(s->fnp, k) * WDWN(k));
      }
      for (int k = 2; k <= k_end + 1; k++) {
        RHS(k) = A3(ph, i, k, j) + A2(s->msfty, i, j) * RHS(k) / (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k));
        if (c->top_lid && k == k_end + 1) RHS(k) = R(0.);
      }
      A3(w, i, 1, j) = A2(s->msfty, i, j) * R(.5) * rdy *
                           ((A2(ht, i, j + 1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                            (A2(ht, i, j) - A2(ht, i, j - 1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j))) +
                       A2(s->msftx, i, j) * R(.5) * rdx *
                           ((A2(ht, i + 1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                            (A2(ht, i, j) - A2(ht, i - 1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j)));
      for (int k = 2; k <= k_end; k++)
        A3(w, i, k, j) = A3(w, i, k, j) + dts * A3(rw_tend, i, k, j) +
                         msft_inv * A3(cqw, i, k, j) *
                             (+R(.5) * dts * K_G * A1(rdn, k) *
                              (A3(c2a, i, k, j) * A1(rdnw, k) / (A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) *
                                   ((R(1.) + epssm) * (RHS(k + 1) - RHS(k)) + (R(1.) - epssm) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j))) -
                               A3(c2a, i, k - 1, j) * A1(rdnw, k - 1) / (A1(c1h, k - 1) * A2(mut, i, j) + A1(c2h, k - 1)) *
                                   ((R(1.) + epssm) * (RHS(k) - RHS(k - 1)) + (R(1.) - epssm) * (A3(ph, i, k, j) - A3(ph, i, k - 1, j))))) +
                         dts * K_G * msft_inv *
                             (A1(rdn, k) * (A3(c2a, i, k, j) * A3(alt, i, k, j) * A3(t_2ave, i, k, j) - A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j)) -
                              (A1(c1f, k) * A2(muave, i, j)));
      {
        int k = k_end + 1;
        A3(w, i, k, j) = A3(w, i, k, j) + dts * A3(rw_tend, i, k, j) +
                         msft_inv * (-R(.5) * dts * K_G / (A1(c1h, k - 1) * A2(mut, i, j) + A1(c2h, k - 1)) * (A1(rdnw, k - 1) * A1(rdnw, k - 1)) * R(2.) *
                                         A3(c2a, i, k - 1, j) * ((R(1.) + epssm) * (RHS(k) - RHS(k - 1)) + (R(1.) - epssm) * (A3(ph, i, k, j) - A3(ph, i, k - 1, j))) -
                                     dts * K_G * (R(2.) * A1(rdnw, k - 1) * A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j) + (A1(c1f, k) * A2(muave, i, j))));
        if (c->top_lid) A3(w, i, k, j) = R(0.);
      }
      for (int k = 2; k <= k_end + 1; k++) A3(w, i, k, j) = (A3(w, i, k, j) - A3(a, i, k, j) * A3(w, i, k - 1, j)) * A3(alpha, i, k, j);
      for (int k = k_end; k >= 2; k--) A3(w, i, k, j) = A3(w, i, k, j) - A3(gamma, i, k, j) * A3(w, i, k + 1, j);
      if (c->damp_opt == 3) {
        for (int k = k_end + 1; k >= 2; k--) {
          wreal htop = (A3(ph_1, i, k_end + 1, j) + A3(phb, i, k_end + 1, j)) / K_G;
          wreal hk = (A3(ph_1, i, k, j) + A3(phb, i, k, j)) / K_G;
          wreal hbot = htop - hdepth;
          wreal dampwt = R(0.);
          if (hk >= hbot) {
            wreal sn = (wreal)sin((double)(R(0.5) * pi * (hk - hbot) / hdepth));
            dampwt = dampmag * sn * sn;
          }
          A3(w, i, k, j) = (A3(w, i, k, j) - dampwt * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)) * A3(w_save, i, k, j)) / (R(1.) + dampwt);
        }
      }
      for (int k = k_end + 1; k >= 2; k--)
        A3(ph, i, k, j) = RHS(k) + A2(s->msfty, i, j) * R(.5) * dts * K_G * (R(1.) + epssm) * A3(w, i, k, j) / (A1(c1f, k) * A2(muts, i, j) + A1(c2f, k));
    }
#undef RHS
#undef WDWN
  free(rhs); free(wdwn);
}

void o_sumflux(ostate *s, const dims_t *d, int iteration, int nsmall) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  wreal *ru = s->u_2, *rv = s->v_2, *ww = s->ww, *u_lin = s->u_save, *v_lin = s->v_save, *ww_lin = s->ww1;
  wreal *ru_m = s->ru_m, *rv_m = s->rv_m, *ww_m = s->ww_m, *c1h = s->c1h, *c2h = s->c2h;
  if (iteration == 1)
    for (int j = jts; j <= jte; j++)
      for (int k = kts; k <= kte; k++)
        for (int i = its; i <= ite; i++) { A3(ru_m, i, k, j) = R(0.); A3(rv_m, i, k, j) = R(0.); A3(ww_m, i, k, j) = R(0.); }
  int mini = IMIN(ide - 1, ite), minj = IMIN(jde - 1, jte), mink = IMIN(kde - 1, kte);
  for (int j = jts; j <= minj; j++)
    for (int k = kts; k <= mink; k++)
      for (int i = its; i <= mini; i++) {
        A3(ru_m, i, k, j) = A3(ru_m, i, k, j) + A3(ru, i, k, j);
        A3(rv_m, i, k, j) = A3(rv_m, i, k, j) + A3(rv, i, k, j);
        A3(ww_m, i, k, j) = A3(ww_m, i, k, j) + A3(ww, i, k, j);
      }
  if (ite > mini)
    for (int j = jts; j <= minj; j++)
      for (int k = kts; k <= mink; k++)
        for (int i = mini + 1; i <= ite; i++) A3(ru_m, i, k, j) = A3(ru_m, i, k, j) + A3(ru, i, k, j);
  if (jte > minj)
    for (int j = minj + 1; j <= jte; j++)
      for (int k = kts; k <= mink; k++)
        for (int i = its; i <= mini; i++) A3(rv_m, i, k, j) = A3(rv_m, i, k, j) + A3(rv, i, k, j);
  if (kte > mink)
    for (int j = jts; j <= minj; j++)
      for (int k = mink + 1; k <= kte; k++)
        for (int i = its; i <= mini; i++) A3(ww_m, i, k, j) = A3(ww_m, i, k, j) + A3(ww, i, k, j);
  if (iteration == nsmall) {
    wreal ns = (wreal)nsmall;
#define RUM(i, k, j) A3(ru_m, i, k, j) = A3(ru_m, i, k, j) / ns + (A1(c1h, k) * A2(s->muu, i, j) + A1(c2h, k)) * A3(u_lin, i, k, j) / A2(s->msfuy, i, j)
#define RVM(i, k, j) A3(rv_m, i, k, j) = A3(rv_m, i, k, j) / ns + (A1(c1h, k) * A2(s->muv, i, j) + A1(c2h, k)) * A3(v_lin, i, k, j) * A2(s->msfvx_inv, i, j)
#define WWM(i, k, j) A3(ww_m, i, k, j) = A3(ww_m, i, k, j) / ns + A3(ww_lin, i, k, j)
    for (int j = jts; j <= minj; j++)
      for (int k = kts; k <= mink; k++)
        for (int i = its; i <= mini; i++) { RUM(i, k, j); RVM(i, k, j); WWM(i, k, j); }
    if (ite > mini)
      for (int j = jts; j <= minj; j++)
        for (int k = kts; k <= mink; k++)
          for (int i = mini + 1; i <= ite; i++) RUM(i, k, j);
    if (jte > minj)
      for (int j = minj + 1; j <= jte; j++)
        for (int k = kts; k <= mink; k++)
          for (int i = its; i <= mini; i++) RVM(i, k, j);
    if (kte > mink)
      for (int j = jts; j <= minj; j++)
        for (int k = mink + 1; k <= kte; k++)
          for (int i = its; i <= mini; i++) WWM(i, k, j);
#undef RUM
#undef RVM
#undef WWM
  }
}

/* module_big_step_utilities_em.F:184-321 with mu = muts, muu = muus, muv = muvs (solve_em.F:1732).
 * The reference's four cases (tile at neither, the first, the last or both domain edges)
 * differ only at i = its when its == ids and at i = ite when ite == ide (likewise in j);
 * every other point takes the centred average. */
void o_calc_mu_uv_1(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  wreal *mu = s->muts, *muu = s->muus, *muv = s->muvs;
  int itf = ite, jtf = IMIN(jte, jde - 1);
  const int ilo = its == ids ? its + 1 : its, ihi = ite == ide ? itf - 1 : itf;
  for (int j = jts; j <= jtf; j++)
    for (int i = ilo; i <= ihi; i++) A2(muu, i, j) = R(0.5) * (A2(mu, i, j) + A2(mu, i - 1, j));
  if (its == ids) {
    int i = its, im = c->periodic_x ? its - 1 : its;
    for (int j = jts; j <= jtf; j++) A2(muu, i, j) = R(0.5) * (A2(mu, i, j) + A2(mu, im, j));
  }
  if (ite == ide) {
    int i = ite, im = c->periodic_x ? ite : ite - 1;
    for (int j = jts; j <= jtf; j++) A2(muu, i, j) = R(0.5) * (A2(mu, i - 1, j) + A2(mu, im, j));
  }
  itf = IMIN(ite, ide - 1); jtf = jte;
  const int jlo = jts == jds ? jts + 1 : jts, jhi = jte == jde ? jtf - 1 : jtf;
  for (int j = jlo; j <= jhi; j++)
    for (int i = its; i <= itf; i++) A2(muv, i, j) = R(0.5) * (A2(mu, i, j) + A2(mu, i, j - 1));
  if (jts == jds) {
    int j = jts, jm = c->periodic_y ? jts - 1 : jts;
    for (int i = its; i <= itf; i++) A2(muv, i, j) = R(0.5) * (A2(mu, i, j) + A2(mu, i, jm));
  }
  if (jte == jde) {
    int j = jte, jm = c->periodic_y ? jte : jte - 1;
    for (int i = its; i <= itf; i++) A2(muv, i, j) = R(0.5) * (A2(mu, i, j - 1) + A2(mu, i, jm));
  }
}
