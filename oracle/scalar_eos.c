/*
 * Oracle: scalar transport and the equation of state.  TEST INFRASTRUCTURE.
 *   rk_scalar_tend    dyn_em/module_em.F:1096-1442 (advect_scalar / advect_scalar_pd, diff_opt 1
 *                     with km_opt 1, sixth-order diffusion)
 *   rk_update_scalar  dyn_em/module_em.F:1587-1799
 *   calc_p_rho_phi    dyn_em/module_big_step_utilities_em.F:953-1223 (non-hydrostatic)
 * Species s->q[m] (m < n_moist are moist species, q[0] = QV); at the solve_em.F
 * call sites (2215, 2309, 2783, 2849) ru/rv/ww are the time-averaged ru_m/rv_m/ww_m,
 * mut is muts, scalar_1 is the *_old array and mu_old/mu_new are mu_1/mu_2.
 */
#include <stdlib.h>
#include "oracle.h"

/* advect_tend is zeroed over the tile (zero_tend, module_em.F:1224) and then filled */
int o_rk_scalar_tend(ostate *s, const cfg_t *c, const dims_t *d, int m, int rk_step, wreal dt_rk) {
  int adv_opt = m < s->n_moist ? c->moist_adv_opt : c->scalar_adv_opt;
  if ((adv_opt != 0 && adv_opt != 1) || c->zadvect_implicit || c->diff_opt < 0 || c->diff_opt > 1 ||
      (c->diff_opt == 1 && c->km_opt != 1) || c->diff_6th_opt < 0 || c->diff_6th_opt > 2)
    return WRFMI_E_BADOPT;
  /* module_em.F:1380-1415: at rk_step 1 with diff_opt 1, horizontal diffusion with xkhh
   * (khdif/prandtl under km_opt 1) and, with no boundary-layer scheme, vertical diffusion
   * with kvdif/prandtl: of qv - qv_base for QV, of the species itself otherwise */
  if (rk_step == 1 && c->diff_opt == 1) {
    const wreal prandtl = R(1.) / R(3.0);
    const wreal khdq = (wreal)c->khdif / prandtl, kvdq = (wreal)c->kvdif / prandtl;
    const size_t n = (size_t)NI_ * NK_ * (size_t)(d->jme - d->jms + 1);
    wreal *xkhh = (wreal *)malloc(n * sizeof(wreal));
    for (size_t q = 0; q < n; q++) xkhh[q] = khdq;
    o_horizontal_diffusion('m', s->q[m], s->q_tend[m], s->muts, s->c1h, s->c2h, c, s->msfux, s->msfuy, s->msfvx,
                           s->msfvx_inv, s->msfvy, s->msftx, s->msfty, khdq, xkhh, s->rdx, s->rdy, d);
    free(xkhh);
    if (c->bl_pbl_physics == 0) {
      if (m == 0 && s->n_moist > 0)
        o_vertical_diffusion_mp(s->q[m], s->q_tend[m], c, s->qv_base, s->c1h, s->c2h, s->alt, s->muts, s->rdn,
                                s->rdnw, kvdq, d);
      else
        o_vertical_diffusion('m', s->q[m], s->q_tend[m], c, s->c1h, s->c2h, s->alt, s->muts, s->rdn, s->rdnw, kvdq, d);
    }
  }
  /* module_em.F:1421-1436: at rk_step 1, 6th-order diffusion of the species into its
   * tendency (the physics tendency array), with rk_scalar_tend's mut = muts and dt_step */
  if (rk_step == 1 && c->diff_6th_opt != 0)
    o_sixth_order_diffusion('m', s->q[m], s->q_tend[m], s->muts, dt_rk, c, s->c1h, s->c2h, c->diff_6th_opt,
                            (wreal)c->diff_6th_factor, s->phb, s->ph_2, s->rdx, s->rdy, s->msftx, s->msfty, s->msfux,
                            s->msfuy, s->msfvx, s->msfvy, d);
  for (int j = d->jts; j <= d->jte; j++)
    for (int k = d->kts; k <= d->kte; k++)
      for (int i = d->its; i <= d->ite; i++) A3(s->advect_tend, i, k, j) = R(0.);
  /* module_em.F:1264-1280: the positive-definite form on the last RK substep, with the
   * time-n values, mub and mu_1 of the limiter; dt_step is the substep's dt */
  if (adv_opt == 1 && rk_step == c->rk_ord) {
    return o_advect_scalar_pd(s->q[m], s->q_old[m], s->advect_tend, NULL, NULL, s->ru_m, s->rv_m, s->ww_m, s->c1h,
                              s->c2h, s->muts, s->mub, s->mu_1, c->time_step, c, 0, s->msfux, s->msfuy, s->msfvx,
                              s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp, s->rdx, s->rdy, s->rdnw, dt_rk, d);
  }
  return o_advect_scalar(s->q[m], s->q[m], s->advect_tend, s->ru_m, s->rv_m, s->ww_m, s->c1h, s->c2h, s->muts,
                         c->time_step, c, s->msfux, s->msfuy, s->msfvx, s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp,
                         s->rdx, s->rdy, s->rdnw, d);
}

void o_rk_update_scalar(ostate *s, const dims_t *d, int m, int rk_step, wreal dt) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  wreal *sc1 = s->q_old[m], *sc2 = s->q[m], *sct = s->q_tend[m];
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      wreal muold = A2(s->mu_1, i, j) + A2(s->mub, i, j);
      wreal munew = A2(s->mu_2, i, j) + A2(s->mub, i, j);
      for (int k = kts; k <= IMIN(kte, kde - 1); k++) {
        /* tendency = advect_tend*msfty (+ sc_tend): periodic/open bounds equal the tile */
        wreal tendency = A3(s->advect_tend, i, k, j) * A2(s->msfty, i, j);
        tendency = tendency + A3(sct, i, k, j);
        if (rk_step == 1) A3(sc1, i, k, j) = A3(sc2, i, k, j);
        A3(sc2, i, k, j) = ((A1(s->c1h, k) * muold + A1(s->c2h, k)) * A3(sc1, i, k, j) + dt * tendency) /
                           (A1(s->c1h, k) * munew + A1(s->c2h, k));
      }
    }
}

/* LOG and ** of calc_p_rho_phi.  Default: the libm routine of the working precision, as a
 * Fortran compiler calls it (logf/powf for default REAL).  With EOS_CORRECTLY_ROUNDED the FP32
 * build instead evaluates them in double and rounds once, which is what the GPU kernels do. */
#if WREAL_IS_DOUBLE
#define EOS_LOG(x) log(x)
#define EOS_POW(x, y) pow(x, y)
#elif defined(EOS_CORRECTLY_ROUNDED)
#define EOS_LOG(x) ((float)log((double)(x)))
#define EOS_POW(x, y) ((float)pow((double)(x), (double)(y)))
#else
#define EOS_LOG(x) logf(x)
#define EOS_POW(x, y) powf(x, y)
#endif

void o_calc_p_rho_phi(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  int itf = IMIN(ite, ide - 1), jtf = IMIN(jte, jde - 1), ktf = IMIN(kte, kde - 1);
  wreal *al = s->al, *alb = s->alb, *mu = s->mu_2, *muts = s->muts, *ph = s->ph_2, *phb = s->phb, *p = s->p, *pb = s->pb, *t = s->t_2;
  wreal ptop = s->p_top;
  for (int j = jts; j <= jtf; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= itf; i++) {
        if (c->hypsometric_opt == 1) {
          A3(al, i, k, j) = -R(1.) / (A1(s->c1h, k) * A2(muts, i, j) + A1(s->c2h, k)) *
                            (A3(alb, i, k, j) * (A1(s->c1h, k) * A2(mu, i, j)) + A1(s->rdnw, k) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)));
        } else {
          wreal pfu = A1(s->c3f, k + 1) * A2(muts, i, j) + A1(s->c4f, k + 1) + ptop;
          wreal pfd = A1(s->c3f, k) * A2(muts, i, j) + A1(s->c4f, k) + ptop;
          wreal phm = A1(s->c3h, k) * A2(muts, i, j) + A1(s->c4h, k) + ptop;
          wreal lg = EOS_LOG(pfd / pfu);
          A3(al, i, k, j) = (A3(ph, i, k + 1, j) - A3(ph, i, k, j) + A3(phb, i, k + 1, j) - A3(phb, i, k, j)) / phm / lg - A3(alb, i, k, j);
        }
      }
  for (int j = jts; j <= jtf; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= itf; i++) {
        wreal temp;
        if (s->n_moist >= 1) {
          if (c->use_theta_m == 1) {
            temp = (K_RD * (K_T0 + A3(t, i, k, j))) / (K_P0 * (A3(al, i, k, j) + A3(alb, i, k, j)));
          } else {
            wreal qvf = R(1.) + K_RVOVRD * A3(s->q[0], i, k, j);
            temp = (K_RD * (K_T0 + A3(t, i, k, j)) * qvf) / (K_P0 * (A3(al, i, k, j) + A3(alb, i, k, j)));
          }
          wreal pw = EOS_POW(temp, K_CPOVCV);
          A3(p, i, k, j) = pw * K_P0 - A3(pb, i, k, j);
        } else {
          A3(p, i, k, j) = K_P0 * EOS_POW((K_RD * (K_T0 + A3(t, i, k, j))) / (K_P0 * (A3(al, i, k, j) + A3(alb, i, k, j))), K_CPOVCV) - A3(pb, i, k, j);
        }
      }
}
