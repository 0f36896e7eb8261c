/*
 * Oracle: lateral boundary (halo) fills and rk_step_prep.  TEST INFRASTRUCTURE.
 *   set_physical_bc3d  share/module_bc.F:651-1113
 *   set_physical_bc2d  share/module_bc.F:202-647
 *   rk_phys_bc_dry_1   dyn_em/module_bc_em.F:960-1043
 *   set_w_surface      dyn_em/module_bc_em.F:1196-1295 (fill_w_flag false)
 *   rk_step_prep       dyn_em/module_em.F:37-186 with calculate_full
 *                      (module_big_step_utilities_em.F:3587-3636), calc_mu_uv (26-180),
 *                      couple_momentum (329-399), calc_ww_cp (640-782), calc_cq (787-906),
 *                      calc_alt (910-949), calc_php (1227-1266)
 */
#include <stdlib.h>
#include "oracle.h"

#define BDYZONE 4 /* share/module_bc.F:33 */

/* One routine for the 3-D and 2-D forms: nk = 0 means 2-D (dat(ims:ime,jms:jme)).
 * The tile is the whole patch (its = ids, ite = ide ...), as in a serial run. */
static void set_bc(wreal *dat, char var_in, const cfg_t *c, const dims_t *d, int is3d) {
  char var = var_in;
  if (var >= 'A' && var <= 'Z') var = (char)(var - 'A' + 'a');
  int open_bc_copy;
  if (is3d)
    open_bc_copy = (var == 'u' || var == 'v' || var == 'w' || var == 't' || var == 'd' ||
                    var == 'e' || var == 'x' || var == 'y' || var == 'f' || var == 'r' || var == 'p');
  else
    open_bc_copy = (var == 'u' || var == 'v' || var == 'w' || var == 't' || var == 'x' ||
                    var == 'y' || var == 'r' || var == 'p');
  int istag = -1, jstag = -1;
  int k_end = IMAX(1, IMIN(d->kde - 1, d->kte));
  if (var == 'u' || var == 'x') istag = 0;
  if (var == 'v' || var == 'y') jstag = 0;
  if (var == 'd') { istag = 0; jstag = 0; }
  if (var == 'e') { istag = 0; if (is3d) k_end = IMIN(d->kde, d->kte); }
  if (var == 'f') { jstag = 0; if (is3d) k_end = IMIN(d->kde, d->kte); }
  if (var == 'w') k_end = IMIN(d->kde, d->kte);
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  int kts = d->kts, kend = k_end;
  if (!is3d) { kts = 0; kend = 0; }
#define DAT(i, k, j) (*(is3d ? &A3(dat, i, k, j) : &A2(dat, i, j)))
  if (c->periodic_x) {
    if (its == ids) {
      for (int j = IMAX(jds, jts - 1); j <= IMIN(jte + 1, jde + jstag); j++)
        for (int k = kts; k <= kend; k++)
          for (int i = 0; i >= -(BDYZONE - 1); i--) DAT(ids + i - 1, k, j) = DAT(ide + i - 1, k, j);
    }
    if (ite == ide) {
      for (int j = IMAX(jds, jts - 1); j <= IMIN(jte + 1, jde + jstag); j++)
        for (int k = kts; k <= kend; k++)
          for (int i = -istag; i <= BDYZONE; i++) DAT(ide + i + istag, k, j) = DAT(ids + i + istag, k, j);
    }
  } else {
    if (c->symmetric_xs && its == ids) {
      for (int j = IMAX(jds, jts - 1); j <= IMIN(jte + 1, jde + jstag); j++)
        for (int k = kts; k <= kend; k++)
          for (int i = 1; i <= BDYZONE; i++) {
            if (istag == -1) DAT(ids - i, k, j) = DAT(ids + i - 1, k, j);
            else if (var == 'u') DAT(ids - i, k, j) = -DAT(ids + i, k, j);
            else DAT(ids - i, k, j) = DAT(ids + i, k, j);
          }
    }
    if (c->symmetric_xe && ite == ide) {
      for (int j = IMAX(jds, jts - 1); j <= IMIN(jte + 1, jde + jstag); j++)
        for (int k = kts; k <= kend; k++)
          for (int i = 1; i <= BDYZONE; i++) {
            if (istag == -1) DAT(ide + i - 1, k, j) = DAT(ide - i, k, j);
            else if (var == 'u') DAT(ide + i, k, j) = -DAT(ide - i, k, j);
            else DAT(ide + i, k, j) = DAT(ide - i, k, j);
          }
    }
    if ((c->open_xs || c->specified || c->nested) && its == ids && open_bc_copy) {
      int j0 = is3d ? jts - BDYZONE : IMAX(jds, jts - 1);
      int j1 = is3d ? IMIN(jte, jde + jstag) + BDYZONE : IMIN(jte + 1, jde + jstag);
      for (int j = j0; j <= j1; j++)
        for (int k = kts; k <= kend; k++) {
          DAT(ids - 1, k, j) = DAT(ids, k, j);
          DAT(ids - 2, k, j) = DAT(ids, k, j);
          DAT(ids - 3, k, j) = DAT(ids, k, j);
        }
    }
    if ((c->open_xe || c->specified || c->nested) && ite == ide && open_bc_copy) {
      if (var != 'u' && var != 'x') {
        int j0 = is3d ? jts - BDYZONE : IMAX(jds, jts - 1);
        int j1 = is3d ? IMIN(jte, jde + jstag) + BDYZONE : IMIN(jte + 1, jde + jstag);
        for (int j = j0; j <= j1; j++)
          for (int k = kts; k <= kend; k++) {
            DAT(ide, k, j) = DAT(ide - 1, k, j);
            DAT(ide + 1, k, j) = DAT(ide - 1, k, j);
            DAT(ide + 2, k, j) = DAT(ide - 1, k, j);
          }
      } else {
        int j0 = is3d ? IMAX(jds, jts - 1) - BDYZONE : IMAX(jds, jts - 1);
        int j1 = is3d ? IMIN(jte + 1, jde + jstag) + BDYZONE : IMIN(jte + 1, jde + jstag);
        for (int j = j0; j <= j1; j++)
          for (int k = kts; k <= kend; k++) {
            DAT(ide + 1, k, j) = DAT(ide, k, j);
            DAT(ide + 2, k, j) = DAT(ide, k, j);
            DAT(ide + 3, k, j) = DAT(ide, k, j);
          }
      }
    }
  }
  int i_start = IMAX(ids, its - 1), i_end = IMIN(ite + 1, ide + istag);
  if (its == ids) i_start = d->ims;
  if (ite == ide) i_end = d->ime;
  if (c->periodic_y) {
    if (jts == jds)
      for (int j = 0; j >= -(BDYZONE - 1); j--)
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) DAT(i, k, jds + j - 1) = DAT(i, k, jde + j - 1);
    if (jte == jde)
      for (int j = -jstag; j <= BDYZONE; j++)
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) DAT(i, k, jde + j + jstag) = DAT(i, k, jds + j + jstag);
  } else {
    if (c->symmetric_ys && jts == jds)
      for (int j = 1; j <= BDYZONE; j++)
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) {
            if (jstag == -1) DAT(i, k, jds - j) = DAT(i, k, jds + j - 1);
            else if (var == 'v') DAT(i, k, jds - j) = -DAT(i, k, jds + j);
            else DAT(i, k, jds - j) = DAT(i, k, jds + j);
          }
    if (c->symmetric_ye && jte == jde)
      for (int j = 1; j <= BDYZONE; j++)
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) {
            if (jstag == -1) DAT(i, k, jde + j - 1) = DAT(i, k, jde - j);
            else if (var == 'v') DAT(i, k, jde + j) = -DAT(i, k, jde - j);
            else DAT(i, k, jde + j) = DAT(i, k, jde - j);
          }
    if ((c->open_ys || c->polar || c->specified || c->nested) && jts == jds && open_bc_copy)
      for (int k = kts; k <= kend; k++)
        for (int i = i_start; i <= i_end; i++) {
          DAT(i, k, jds - 1) = DAT(i, k, jds);
          DAT(i, k, jds - 2) = DAT(i, k, jds);
          DAT(i, k, jds - 3) = DAT(i, k, jds);
        }
    if ((c->open_ye || c->polar || c->specified || c->nested) && jte == jde && open_bc_copy) {
      if (var != 'v' && var != 'y') {
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) {
            DAT(i, k, jde) = DAT(i, k, jde - 1);
            DAT(i, k, jde + 1) = DAT(i, k, jde - 1);
            DAT(i, k, jde + 2) = DAT(i, k, jde - 1);
          }
      } else {
        for (int k = kts; k <= kend; k++)
          for (int i = i_start; i <= i_end; i++) {
            DAT(i, k, jde + 1) = DAT(i, k, jde);
            DAT(i, k, jde + 2) = DAT(i, k, jde);
            DAT(i, k, jde + 3) = DAT(i, k, jde);
          }
      }
    }
  }
#undef DAT
}

void o_set_physical_bc3d(wreal *dat, char var, const cfg_t *c, const dims_t *d) { set_bc(dat, var, c, d, 1); }
void o_set_physical_bc2d(wreal *dat, char var, const cfg_t *c, const dims_t *d) { set_bc(dat, var, c, d, 0); }

/* The kinematic lower boundary condition w(1) = msfty*v.dz/dy + msftx*u.dz/dx, with u and v
 * extrapolated to the surface by cf1..cf3 and the height differences of ht taken on both
 * sides of the point; a neighbour across a non-periodic edge is the point itself
 * (dyn_em/module_bc_em.F:1196-1295, fill_w_flag false: only level 1 is written). */
void o_set_w_surface(wreal *w, const ostate *s, const cfg_t *c, const dims_t *d) {
  const wreal *ht = s->ht, *u = s->u_2, *v = s->v_2;
  const wreal cf1 = s->cf1, cf2 = s->cf2, cf3 = s->cf3, rdx = s->rdx, rdy = s->rdy;
  int jm1_limit = d->jds, jp1_limit = d->jde - 1, im1_limit = d->ids, ip1_limit = d->ide - 1;
  if (c->periodic_x) { im1_limit = d->ids - 1; ip1_limit = d->ide; }
  if (c->periodic_y) { jm1_limit = d->jds - 1; jp1_limit = d->jde; }
  for (int j = d->jts; j <= IMIN(d->jte, d->jde - 1); j++) {
    const int jm1 = IMAX(j - 1, jm1_limit), jp1 = IMIN(j + 1, jp1_limit);
    for (int i = d->its; i <= IMIN(d->ite, d->ide - 1); i++) {
      const int im1 = IMAX(i - 1, im1_limit), ip1 = IMIN(i + 1, ip1_limit);
      A3(w, i, 1, j) = A2(s->msfty, i, j) * R(.5) * rdy *
                           ((A2(ht, i, jp1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                            (A2(ht, i, j) - A2(ht, i, jm1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j))) +
                       A2(s->msftx, i, j) * R(.5) * rdx *
                           ((A2(ht, ip1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                            (A2(ht, i, j) - A2(ht, im1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j)));
    }
  }
}

/* dyn_em/module_bc_em.F:960-1043 (called with ru, rv, rw, ww, muu, muv, mut, php, alt, p) */
void o_rk_phys_bc_dry_1(ostate *s, const cfg_t *c, const dims_t *d) {
  o_set_physical_bc3d(s->ru, 'u', c, d);
  o_set_physical_bc3d(s->rv, 'v', c, d);
  o_set_physical_bc3d(s->rw, 'w', c, d);
  o_set_physical_bc3d(s->ww, 'w', c, d);
  o_set_physical_bc3d(s->php, 'w', c, d);
  o_set_physical_bc3d(s->alt, 't', c, d);
  o_set_physical_bc3d(s->p, 'p', c, d);
  o_set_physical_bc2d(s->muu, 'u', c, d);
  o_set_physical_bc2d(s->muv, 'v', c, d);
  o_set_physical_bc2d(s->mut, 't', c, d);
}

/* module_big_step_utilities_em.F:26-180 */
/* calc_mu_uv (module_big_step_utilities_em.F:26-180): as calc_mu_uv_1 with mub added, in
 * the reference's four tile cases (they differ only at an edge tile's first or last point) */
static void calc_mu_uv(const cfg_t *c, const wreal *mu, const wreal *mub, wreal *muu, wreal *muv,
                       const dims_t *d) {
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde;
  int itf = ite, jtf = IMIN(jte, jde - 1);
#define MUUF(i, im, j) (R(0.5) * (((A2(mu, i, j) + A2(mu, im, j)) + A2(mub, i, j)) + A2(mub, im, j)))
  const int ilo = its == ids ? its + 1 : its, ihi = ite == ide ? itf - 1 : itf;
  for (int j = jts; j <= jtf; j++)
    for (int i = ilo; i <= ihi; i++) A2(muu, i, j) = MUUF(i, i - 1, j);
  if (its == ids) {
    int i = its, im = its;
    if (c->periodic_x) im = its - 1;
    for (int j = jts; j <= jtf; j++) A2(muu, i, j) = MUUF(i, im, j);
  }
  if (ite == ide) {
    int i = ite, im = ite - 1;
    if (c->periodic_x) im = ite;
    /* MUU(i,j) = 0.5*(MU(i-1,j)+MU(im,j)+MUB(i-1,j)+MUB(im,j)) */
    for (int j = jts; j <= jtf; j++)
      A2(muu, i, j) = R(0.5) * (((A2(mu, i - 1, j) + A2(mu, im, j)) + A2(mub, i - 1, j)) + A2(mub, im, j));
  }
#undef MUUF
  itf = IMIN(ite, ide - 1); jtf = jte;
  const int jlo = jts == jds ? jts + 1 : jts, jhi = jte == jde ? jtf - 1 : jtf;
  for (int j = jlo; j <= jhi; j++)
    for (int i = its; i <= itf; i++)
      A2(muv, i, j) = R(0.5) * (((A2(mu, i, j) + A2(mu, i, j - 1)) + A2(mub, i, j)) + A2(mub, i, j - 1));
  if (jts == jds) {
    int j = jts, jm = jts;
    if (c->periodic_y) jm = jts - 1;
    for (int i = its; i <= itf; i++)
      A2(muv, i, j) = R(0.5) * (((A2(mu, i, j) + A2(mu, i, jm)) + A2(mub, i, j)) + A2(mub, i, jm));
  }
  if (jte == jde) {
    int j = jte, jm = jte - 1;
    if (c->periodic_y) jm = jte;
    for (int i = its; i <= itf; i++)
      A2(muv, i, j) = R(0.5) * (((A2(mu, i, j - 1) + A2(mu, i, jm)) + A2(mub, i, j - 1)) + A2(mub, i, jm));
  }
}

void o_rk_step_prep(ostate *s, const cfg_t *c, const dims_t *d, int rk_step) {
  (void)rk_step;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  /* calculate_full (mut = mub + mu), module_big_step_utilities_em.F:3587-3636 with
   * kds..kde = 1..2: it covers its-1..min(ite,ide-1), jts-1..min(jte,jde-1) */
  for (int j = jts - 1; j <= IMIN(jte, jde - 1); j++)
    for (int i = its - 1; i <= IMIN(ite, ide - 1); i++) A2(s->mut, i, j) = A2(s->mub, i, j) + A2(s->mu_2, i, j);
  calc_mu_uv(c, s->mu_2, s->mub, s->muu, s->muv, d);
  /* couple_momentum */
  int ktf = IMIN(kte, kde - 1);
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= ite; i++)
        A3(s->ru, i, k, j) = A3(s->u_2, i, k, j) * (A1(s->c1h, k) * A2(s->muu, i, j) + A1(s->c2h, k)) / A2(s->msfuy, i, j);
  for (int j = jts; j <= jte; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++)
        A3(s->rv, i, k, j) = A3(s->v_2, i, k, j) * (A1(s->c1h, k) * A2(s->muv, i, j) + A1(s->c2h, k)) * A2(s->msfvx_inv, i, j);
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= kte; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++)
        A3(s->rw, i, k, j) = A3(s->w_2, i, k, j) * (A1(s->c1f, k) * A2(s->mut, i, j) + A1(s->c2f, k)) / A2(s->msfty, i, j);
  /* calc_ww_cp: local muu/muv of its own (full column mass at u/v points) */
  {
    int jtf = IMIN(jte, jde - 1), itf = IMIN(ite, ide - 1);
    const wreal *mup = s->mu_2, *mub = s->mub;
    int ni = NI_;
    wreal *lmuu = (wreal *)calloc((size_t)ni * (size_t)(d->jme - d->jms + 1), sizeof(wreal));
    wreal *lmuv = (wreal *)calloc((size_t)ni * (size_t)(d->jme - d->jms + 1), sizeof(wreal));
    for (int j = jts; j <= jtf; j++)
      for (int i = its; i <= IMIN(ite + 1, ide); i++)
        A2(lmuu, i, j) = R(0.5) * (((A2(mup, i, j) + A2(mub, i, j)) + A2(mup, i - 1, j)) + A2(mub, i - 1, j));
    for (int j = jts; j <= IMIN(jte + 1, jde); j++)
      for (int i = its; i <= itf; i++)
        A2(lmuv, i, j) = R(0.5) * (((A2(mup, i, j) + A2(mub, i, j)) + A2(mup, i, j - 1)) + A2(mub, i, j - 1));
    int nk = NK_;
    wreal *divv = (wreal *)malloc(sizeof(wreal) * (size_t)nk);
    for (int j = jts; j <= jtf; j++) {
      for (int i = its; i <= ite; i++) {
        A3(s->ww, i, 1, j) = R(0.);
        A3(s->ww, i, kte, j) = R(0.);
      }
      for (int i = its; i <= itf; i++) {
        wreal dmdt = R(0.);
        for (int k = kts; k <= ktf; k++) {
          wreal c1 = A1(s->c1h, k), c2 = A1(s->c2h, k);
          wreal tx = (c1 * A2(lmuu, i + 1, j) + c2) * A3(s->u_2, i + 1, k, j) / A2(s->msfuy, i + 1, j) -
                     (c1 * A2(lmuu, i, j) + c2) * A3(s->u_2, i, k, j) / A2(s->msfuy, i, j);
          wreal ty = (c1 * A2(lmuv, i, j + 1) + c2) * A3(s->v_2, i, k, j + 1) * A2(s->msfvx_inv, i, j + 1) -
                     (c1 * A2(lmuv, i, j) + c2) * A3(s->v_2, i, k, j) * A2(s->msfvx_inv, i, j);
          divv[k - d->kms] = A2(s->msftx, i, j) * A1(s->dnw, k) * (s->rdx * tx + s->rdy * ty);
          dmdt = dmdt + divv[k - d->kms];
        }
        for (int k = 2; k <= ktf; k++)
          A3(s->ww, i, k, j) = A3(s->ww, i, k - 1, j) - A1(s->dnw, k - 1) * A1(s->c1h, k - 1) * dmdt - divv[k - 1 - d->kms];
      }
    }
    free(divv); free(lmuu); free(lmuv);
  }
  /* calc_cq: moist species PARAM_FIRST_SCALAR..n_moist are s->q[0..n_moist-1] */
  if (s->n_moist >= 1) {
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= ite; i++) {
          wreal qtot = R(0.);
          for (int m = 0; m < s->n_moist; m++) qtot = qtot + A3(s->q[m], i, k, j) + A3(s->q[m], i - 1, k, j);
          A3(s->cqu, i, k, j) = R(1.) / (R(1.) + R(0.5) * qtot);
        }
    for (int j = jts; j <= jte; j++)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= IMIN(ite, ide - 1); i++) {
          wreal qtot = R(0.);
          for (int m = 0; m < s->n_moist; m++) qtot = qtot + A3(s->q[m], i, k, j) + A3(s->q[m], i, k, j - 1);
          A3(s->cqv, i, k, j) = R(1.) / (R(1.) + R(0.5) * qtot);
        }
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts + 1; k <= ktf; k++)
        for (int i = its; i <= IMIN(ite, ide - 1); i++) {
          wreal qtot = R(0.);
          for (int m = 0; m < s->n_moist; m++) qtot = qtot + A3(s->q[m], i, k, j) + A3(s->q[m], i, k - 1, j);
          A3(s->cqw, i, k, j) = R(0.5) * qtot;
        }
  } else {
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= ite; i++) A3(s->cqu, i, k, j) = R(1.);
    for (int j = jts; j <= jte; j++)
      for (int k = kts; k <= ktf; k++)
        for (int i = its; i <= IMIN(ite, ide - 1); i++) A3(s->cqv, i, k, j) = R(1.);
    for (int j = jts; j <= IMIN(jte, jde - 1); j++)
      for (int k = kts + 1; k <= ktf; k++)
        for (int i = its; i <= IMIN(ite, ide - 1); i++) A3(s->cqw, i, k, j) = R(0.);
  }
  /* calc_alt, calc_php */
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        A3(s->alt, i, k, j) = A3(s->al, i, k, j) + A3(s->alb, i, k, j);
        A3(s->php, i, k, j) = R(0.5) * (((A3(s->phb, i, k, j) + A3(s->phb, i, k + 1, j)) + A3(s->ph_2, i, k, j)) +
                                        A3(s->ph_2, i, k + 1, j));
      }
}
