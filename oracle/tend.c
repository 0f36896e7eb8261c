/*
 * Oracle: large-step (RK) tendencies.  TEST INFRASTRUCTURE.
 *   rk_tendency                  dyn_em/module_em.F:190-955 (physics-off path)
 *   rhs_ph                       dyn_em/module_big_step_utilities_em.F:1365-2178
 *   horizontal_pressure_gradient dyn_em/module_big_step_utilities_em.F:2183-2415
 *   pg_buoy_w                    dyn_em/module_big_step_utilities_em.F:2419-2499
 *   w_damp                       dyn_em/module_big_step_utilities_em.F:2503-2711
 *   coriolis                     dyn_em/module_big_step_utilities_em.F:3640-3850
 *   curvature                    dyn_em/module_big_step_utilities_em.F:4175-4469
 *   zero_tend / zero_tend2d      dyn_em/module_big_step_utilities_em.F:4573-4646
 *   rk_addtend_dry               dyn_em/module_em.F:959-1092
 */
#include <stdlib.h>
#include "oracle.h"

static void zero3(wreal *a, const dims_t *d) {
  for (int j = d->jts; j <= d->jte; j++)
    for (int k = d->kts; k <= d->kte; k++)
      for (int i = d->its; i <= d->ite; i++) A3(a, i, k, j) = R(0.);
}
static void zero2(wreal *a, const dims_t *d) {
  for (int j = d->jts; j <= d->jte; j++)
    for (int i = d->its; i <= d->ite; i++) A2(a, i, j) = R(0.);
}

/* module_big_step_utilities_em.F:1365-2178; ph_old == ph (module_em.F:668) */
static void rhs_ph(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kte = d->kte;
  const int specified = c->specified || c->nested;
  wreal *pt = s->ph_tend;
  const wreal *u = s->u_2, *v = s->v_2, *ww = s->ww, *ph = s->ph_2, *phb = s->phb, *w = s->w_2;
  const wreal *mut = s->mut, *muuf = s->muu, *muvf = s->muv;
  int itf = IMIN(ite, ide - 1), jtf = IMIN(jte, jde - 1);
  wreal *wdwn = (wreal *)calloc((size_t)NK_ + 1, sizeof(wreal));
#define WD(k) wdwn[(k) - d->kms]
  for (int j = jts; j <= jtf; j++)
    for (int i = its; i <= itf; i++) {
      if (c->phi_adv_z == 2) {
        for (int k = 2; k <= kte; k++)
          WD(k) = A1(s->rdnw, k - 1) * (A3(ph, i, k, j) - A3(ph, i, k - 1, j) + A3(phb, i, k, j) - A3(phb, i, k - 1, j));
        for (int k = 2; k <= kte - 1; k++)
          A3(pt, i, k, j) = A3(pt, i, k, j) - A3(ww, i, k, j) * (A1(s->fnm, k) * WD(k + 1) + A1(s->fnp, k) * WD(k));
      } else {
        for (int k = 2; k <= kte; k++)
          WD(k) = R(.5) * (A3(ww, i, k, j) + A3(ww, i, k - 1, j)) * A1(s->rdnw, k - 1) *
                  (A3(ph, i, k, j) - A3(ph, i, k - 1, j) + A3(phb, i, k, j) - A3(phb, i, k - 1, j));
        for (int k = 2; k <= kte - 1; k++)
          A3(pt, i, k, j) = A3(pt, i, k, j) - (A1(s->fnm, k) * WD(k + 1) + A1(s->fnp, k) * WD(k));
      }
    }
#undef WD
  free(wdwn);
  if (c->non_hydrostatic)
    for (int j = jts; j <= jtf; j++) {
      for (int i = its; i <= itf; i++) A3(pt, i, kde, j) = R(0.);
      for (int k = 2; k <= kte; k++)
        for (int i = its; i <= itf; i++)
          A3(pt, i, k, j) = A3(pt, i, k, j) + (A1(s->c1f, k) * A2(mut, i, j) + A1(s->c2f, k)) * K_G * A3(w, i, k, j) / A2(s->msfty, i, j);
    }
  const wreal cfn = s->cfn, cfn1 = s->cfn1, rdx = s->rdx, rdy = s->rdy;
  /* y: face mass fluxes at j and j+1; kk = kte uses the extrapolated (cfn, cfn1) velocity */
#define VFACE(i, k, jj) ((k) < kte ? (A1(s->c1f, k) * A2(muvf, i, jj) + A1(s->c2f, k)) * (A3(v, i, k, jj) + A3(v, i, (k) - 1, jj)) * A2(s->msfvy, i, jj) \
                                   : (A1(s->c1f, k) * A2(muvf, i, jj) + A1(s->c2f, k)) * (cfn * A3(v, i, (k) - 1, jj) + cfn1 * A3(v, i, (k) - 2, jj)) * A2(s->msfvy, i, jj))
#define UFACE(ii, k, j) ((k) < kte ? (A1(s->c1f, k) * A2(muuf, ii, j) + A1(s->c2f, k)) * (A3(u, ii, k, j) + A3(u, ii, (k) - 1, j)) * A2(s->msfux, ii, j) \
                                   : (A1(s->c1f, k) * A2(muuf, ii, j) + A1(s->c2f, k)) * (cfn * A3(u, ii, (k) - 1, j) + cfn1 * A3(u, ii, (k) - 2, j)) * A2(s->msfux, ii, j))
#define CY(i, k, j) ((k) < kte ? (R(0.25) * rdy / A2(s->msfty, i, j)) : (R(0.5) * rdy / A2(s->msfty, i, j)))
#define CX(i, k, j) ((k) < kte ? (R(0.25) * rdx / A2(s->msfty, i, j)) : (R(0.5) * rdx / A2(s->msfty, i, j)))
  if (c->h_sca_adv_order <= 6 && c->h_sca_adv_order > 4) {
    int j_start = jts;
    itf = IMIN(ite, ide - 1); jtf = IMIN(jte, jde - 1);
    if (c->open_ys || specified) j_start = IMAX(jts, jds + 3);
    if (c->open_ye || specified) jtf = IMIN(jtf, jde - 4);
    for (int j = j_start; j <= jtf; j++)
      for (int k = 2; k <= kte; k++)
        for (int i = its; i <= itf; i++) {
          wreal S = R(45.) * (A3(ph, i, k, j + 1) - A3(ph, i, k, j - 1)) - R(9.) * (A3(ph, i, k, j + 2) - A3(ph, i, k, j - 2)) +
                    (A3(ph, i, k, j + 3) - A3(ph, i, k, j - 3)) + R(45.) * (A3(phb, i, k, j + 1) - A3(phb, i, k, j - 1)) -
                    R(9.) * (A3(phb, i, k, j + 2) - A3(phb, i, k, j - 2)) + (A3(phb, i, k, j + 3) - A3(phb, i, k, j - 3));
          A3(pt, i, k, j) = A3(pt, i, k, j) - CY(i, k, j) * ((VFACE(i, k, j + 1) + VFACE(i, k, j)) * (R(1.) / R(60.)) * S);
        }
    int jl[2] = {jds + 2, jde - 3};
    int jon[2] = {(c->open_ys || specified) && jts <= jds + 2 && jds + 2 <= jte, (c->open_ye || specified) && jts <= jde - 3 && jde - 3 <= jte};
    for (int b = 0; b < 2; b++)
      if (jon[b]) {
        int j = jl[b];
        for (int k = 2; k <= kte; k++)
          for (int i = its; i <= itf; i++) {
            wreal S = R(8.) * (A3(ph, i, k, j + 1) - A3(ph, i, k, j - 1)) - (A3(ph, i, k, j + 2) - A3(ph, i, k, j - 2)) +
                      R(8.) * (A3(phb, i, k, j + 1) - A3(phb, i, k, j - 1)) - (A3(phb, i, k, j + 2) - A3(phb, i, k, j - 2));
            A3(pt, i, k, j) = A3(pt, i, k, j) - CY(i, k, j) * ((VFACE(i, k, j + 1) + VFACE(i, k, j)) * (R(1.) / R(12.)) * S);
          }
      }
    int j2[2] = {jds + 1, jde - 2};
    int j2on[2] = {(c->open_ys || specified) && jts <= jds + 1 && jds + 1 <= jte, (c->open_ye || specified) && jts <= jde - 2 && jde - 2 <= jte};
    for (int b = 0; b < 2; b++)
      if (j2on[b]) {
        int j = j2[b];
        for (int k = 2; k <= kte; k++)
          for (int i = its; i <= itf; i++) {
            wreal dA = A3(phb, i, k, j + 1) - A3(phb, i, k, j) + A3(ph, i, k, j + 1) - A3(ph, i, k, j);
            wreal dB = A3(phb, i, k, j) - A3(phb, i, k, j - 1) + A3(ph, i, k, j) - A3(ph, i, k, j - 1);
            A3(pt, i, k, j) = A3(pt, i, k, j) - CY(i, k, j) * (VFACE(i, k, j + 1) * dA + VFACE(i, k, j) * dB);
          }
      }
    int i_start = its;
    itf = IMIN(ite, ide - 1); jtf = IMIN(jte, jde - 1);
    if (c->open_xs || specified) i_start = IMAX(its, ids + 3);
    if (c->open_xe || specified) itf = IMIN(itf, ide - 4);
    for (int j = jts; j <= jtf; j++)
      for (int k = 2; k <= kte; k++)
        for (int i = i_start; i <= itf; i++) {
          wreal S = R(45.) * (A3(ph, i + 1, k, j) - A3(ph, i - 1, k, j)) - R(9.) * (A3(ph, i + 2, k, j) - A3(ph, i - 2, k, j)) +
                    (A3(ph, i + 3, k, j) - A3(ph, i - 3, k, j)) + R(45.) * (A3(phb, i + 1, k, j) - A3(phb, i - 1, k, j)) -
                    R(9.) * (A3(phb, i + 2, k, j) - A3(phb, i - 2, k, j)) + (A3(phb, i + 3, k, j) - A3(phb, i - 3, k, j));
          A3(pt, i, k, j) = A3(pt, i, k, j) - CX(i, k, j) * ((UFACE(i + 1, k, j) + UFACE(i, k, j)) * (R(1.) / R(60.)) * S);
        }
    int il[2] = {ids + 2, ide - 3};
    int ion[2] = {c->open_xs && its <= ids + 2 && ids + 2 <= ite, c->open_xe && its <= ide - 3 && ide - 3 <= ite};
    for (int b = 0; b < 2; b++)
      if (ion[b]) {
        int i = il[b];
        for (int j = jts; j <= jtf; j++)
          for (int k = 2; k <= kte; k++) {
            wreal S = R(8.) * (A3(ph, i + 1, k, j) - A3(ph, i - 1, k, j)) - (A3(ph, i + 2, k, j) - A3(ph, i - 2, k, j)) +
                      R(8.) * (A3(phb, i + 1, k, j) - A3(phb, i - 1, k, j)) - (A3(phb, i + 2, k, j) - A3(phb, i - 2, k, j));
            A3(pt, i, k, j) = A3(pt, i, k, j) - CX(i, k, j) * ((UFACE(i + 1, k, j) + UFACE(i, k, j)) * (R(1.) / R(12.)) * S);
          }
      }
    int i2[2] = {ids + 1, ide - 2};
    int i2on[2] = {(c->open_xs || specified) && its <= ids + 1 && ids + 1 <= ite, (c->open_xe || specified) && its <= ide - 2 && ide - 2 <= ite};
    for (int b = 0; b < 2; b++)
      if (i2on[b]) {
        int i = i2[b];
        for (int j = jts; j <= jtf; j++)
          for (int k = 2; k <= kte; k++) {
            wreal dA = A3(phb, i + 1, k, j) - A3(phb, i, k, j) + A3(ph, i + 1, k, j) - A3(ph, i, k, j);
            wreal dB = A3(phb, i, k, j) - A3(phb, i - 1, k, j) + A3(ph, i, k, j) - A3(ph, i - 1, k, j);
            A3(pt, i, k, j) = A3(pt, i, k, j) - CX(i, k, j) * (UFACE(i + 1, k, j) * dA + UFACE(i, k, j) * dB);
          }
      }
  }
#undef VFACE
#undef UFACE
#undef CX
#undef CY
  /* lateral open boundary conditions (ph_old == ph) */
  itf = IMIN(ite, ide - 1);
  if (c->open_ys && jts == jds) {
    int j = jts;
    for (int k = 2; k <= kde; k++) {
      int kz = IMIN(k, kde - 1);
      for (int i = its; i <= itf; i++) {
        wreal vb = R(.5) * (A1(s->fnm, kz) * (A3(v, i, kz, j + 1) + A3(v, i, kz, j)) + A1(s->fnp, kz) * (A3(v, i, kz - 1, j + 1) + A3(v, i, kz - 1, j)));
        wreal vl = vb < R(0.) ? vb : R(0.);
        A3(pt, i, k, j) = A3(pt, i, k, j) - rdy * (A1(s->c1f, k) * A2(mut, i, j) + A1(s->c2f, k)) * (vl * (A3(ph, i, k, j + 1) - A3(ph, i, k, j)));
      }
    }
  }
  if (c->open_ye && jte == jde) {
    int j = jte - 1;
    for (int k = 2; k <= kde; k++) {
      int kz = IMIN(k, kde - 1);
      for (int i = its; i <= itf; i++) {
        wreal vb = R(.5) * (A1(s->fnm, kz) * (A3(v, i, kz, j + 1) + A3(v, i, kz, j)) + A1(s->fnp, kz) * (A3(v, i, kz - 1, j + 1) + A3(v, i, kz - 1, j)));
        wreal vr = vb > R(0.) ? vb : R(0.);
        A3(pt, i, k, j) = A3(pt, i, k, j) - rdy * (A1(s->c1f, k) * A2(mut, i, j) + A1(s->c2f, k)) * (vr * (A3(ph, i, k, j) - A3(ph, i, k, j - 1)));
      }
    }
  }
  jtf = IMIN(jte, jde - 1);
  if (c->open_xs && its == ids) {
    int i = its;
    for (int j = jts; j <= jtf; j++)
      for (int k = 2; k <= kde; k++) {
        int kz = k; /* the reference uses kz = k also at k = kde on this side */
        wreal ub = R(.5) * (A1(s->fnm, kz) * (A3(u, i + 1, kz, j) + A3(u, i, kz, j)) + A1(s->fnp, kz) * (A3(u, i + 1, kz - 1, j) + A3(u, i, kz - 1, j)));
        wreal ul = ub < R(0.) ? ub : R(0.);
        A3(pt, i, k, j) = A3(pt, i, k, j) - (A2(s->msftx, i, j) / A2(s->msfty, i, j)) * rdx * (A1(s->c1f, k) * A2(mut, i, j) + A1(s->c2f, k)) *
                                                (ul * (A3(ph, i + 1, k, j) - A3(ph, i, k, j)));
      }
  }
  if (c->open_xe && ite == ide) {
    int i = ite - 1;
    for (int j = jts; j <= jtf; j++)
      for (int k = 2; k <= kde; k++) {
        int kz = k < kde ? k : k - 1;
        wreal ub = R(.5) * (A1(s->fnm, kz) * (A3(u, i + 1, kz, j) + A3(u, i, kz, j)) + A1(s->fnp, kz) * (A3(u, i + 1, kz - 1, j) + A3(u, i, kz - 1, j)));
        wreal ur = ub > R(0.) ? ub : R(0.);
        A3(pt, i, k, j) = A3(pt, i, k, j) - (A2(s->msftx, i, j) / A2(s->msfty, i, j)) * rdx * (A1(s->c1f, k) * A2(mut, i, j) + A1(s->c2f, k)) *
                                                (ur * (A3(ph, i, k, j) - A3(ph, i - 1, k, j)));
      }
  }
}

/* module_big_step_utilities_em.F:2183-2415 (mu = mu_2, c1 = c1h) */
static void horizontal_pressure_gradient(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kte = d->kte;
  const int specified = c->specified || c->nested;
  const wreal *ph = s->ph_2, *alt = s->alt, *p = s->p, *pb = s->pb, *al = s->al, *php = s->php, *mu = s->mu_2;
  const wreal rdx = s->rdx, rdy = s->rdy, cf1 = s->cf1, cf2 = s->cf2, cf3 = s->cf3, cfn = s->cfn, cfn1 = s->cfn1;
  int ktf = IMIN(kte, kde - 1);
  wreal *dpn = (wreal *)calloc((size_t)NK_ + 1, sizeof(wreal));
#define DPN(k) dpn[(k) - d->kms]
  {
    int itf = IMIN(ite, ide - 1), jtf = jte, j_start = jts;
    if ((c->open_ys || specified || c->polar) && jts == jds) j_start = jts + 1;
    if ((c->open_ye || specified || c->polar) && jte == jde) jtf = jtf - 1;
    for (int j = j_start; j <= jtf; j++)
      for (int i = its; i <= itf; i++) {
        if (c->non_hydrostatic) {
          DPN(1) = R(.5) * (cf1 * (A3(p, i, 1, j - 1) + A3(p, i, 1, j)) + cf2 * (A3(p, i, 2, j - 1) + A3(p, i, 2, j)) +
                            cf3 * (A3(p, i, 3, j - 1) + A3(p, i, 3, j)));
          DPN(kde) = R(0.);
          if (c->top_lid)
            DPN(kde) = R(.5) * (cfn * (A3(p, i, kde - 1, j - 1) + A3(p, i, kde - 1, j)) + cfn1 * (A3(p, i, kde - 2, j - 1) + A3(p, i, kde - 2, j)));
          for (int k = 2; k <= ktf; k++)
            DPN(k) = R(.5) * (A1(s->fnm, k) * (A3(p, i, k, j - 1) + A3(p, i, k, j)) + A1(s->fnp, k) * (A3(p, i, k - 1, j - 1) + A3(p, i, k - 1, j)));
        }
        for (int k = 1; k <= ktf; k++) {
          wreal dpy = (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * R(.5) * rdy * (A1(s->c1h, k) * A2(s->muv, i, j) + A1(s->c2h, k)) *
                      ((A3(ph, i, k + 1, j) - A3(ph, i, k + 1, j - 1) + A3(ph, i, k, j) - A3(ph, i, k, j - 1)) +
                       (A3(alt, i, k, j) + A3(alt, i, k, j - 1)) * (A3(p, i, k, j) - A3(p, i, k, j - 1)) +
                       (A3(al, i, k, j) + A3(al, i, k, j - 1)) * (A3(pb, i, k, j) - A3(pb, i, k, j - 1)));
          if (c->non_hydrostatic) {
            dpy = dpy + (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * rdy * (A3(php, i, k, j) - A3(php, i, k, j - 1)) *
                            (A1(s->rdnw, k) * (DPN(k + 1) - DPN(k)) - R(.5) * ((A1(s->c1h, k) * A2(mu, i, j - 1)) + (A1(s->c1h, k) * A2(mu, i, j))));
            A3(s->rv_tend, i, k, j) = A3(s->rv_tend, i, k, j) - A3(s->cqv, i, k, j) * dpy;
          } else {
            A3(s->rv_tend, i, k, j) = A3(s->rv_tend, i, k, j) - dpy;
          }
        }
      }
  }
  {
    int itf = ite, jtf = IMIN(jte, jde - 1), i_start = its;
    if ((c->open_xs || specified || c->nested) && its == d->ids) i_start = its + 1;
    if ((c->open_xe || specified || c->nested) && ite == ide) itf = itf - 1;
    if (c->periodic_x) { i_start = its; itf = ite; }
    for (int j = jts; j <= jtf; j++)
      for (int i = i_start; i <= itf; i++) {
        if (c->non_hydrostatic) {
          DPN(1) = R(.5) * (cf1 * (A3(p, i - 1, 1, j) + A3(p, i, 1, j)) + cf2 * (A3(p, i - 1, 2, j) + A3(p, i, 2, j)) +
                            cf3 * (A3(p, i - 1, 3, j) + A3(p, i, 3, j)));
          DPN(kde) = R(0.);
          if (c->top_lid)
            DPN(kde) = R(.5) * (cfn * (A3(p, i - 1, kde - 1, j) + A3(p, i, kde - 1, j)) + cfn1 * (A3(p, i - 1, kde - 2, j) + A3(p, i, kde - 2, j)));
          for (int k = 2; k <= ktf; k++)
            DPN(k) = R(.5) * (A1(s->fnm, k) * (A3(p, i - 1, k, j) + A3(p, i, k, j)) + A1(s->fnp, k) * (A3(p, i - 1, k - 1, j) + A3(p, i, k - 1, j)));
        }
        for (int k = 1; k <= ktf; k++) {
          wreal dpx = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * R(.5) * rdx * (A1(s->c1h, k) * A2(s->muu, i, j) + A1(s->c2h, k)) *
                      ((A3(ph, i, k + 1, j) - A3(ph, i - 1, k + 1, j) + A3(ph, i, k, j) - A3(ph, i - 1, k, j)) +
                       (A3(alt, i, k, j) + A3(alt, i - 1, k, j)) * (A3(p, i, k, j) - A3(p, i - 1, k, j)) +
                       (A3(al, i, k, j) + A3(al, i - 1, k, j)) * (A3(pb, i, k, j) - A3(pb, i - 1, k, j)));
          if (c->non_hydrostatic) {
            dpx = dpx + (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * rdx * (A3(php, i, k, j) - A3(php, i - 1, k, j)) *
                            (A1(s->rdnw, k) * (DPN(k + 1) - DPN(k)) - R(.5) * ((A1(s->c1h, k) * A2(mu, i - 1, j)) + (A1(s->c1h, k) * A2(mu, i, j))));
            A3(s->ru_tend, i, k, j) = A3(s->ru_tend, i, k, j) - A3(s->cqu, i, k, j) * dpx;
          } else {
            A3(s->ru_tend, i, k, j) = A3(s->ru_tend, i, k, j) - dpx;
          }
        }
      }
  }
#undef DPN
  free(dpn);
}

/* module_big_step_utilities_em.F:2419-2499 (muf = mu_2, mubf = mub); overwrites cqw(2:kde-1) with 1/(1+cqw) */
static void pg_buoy_w(ostate *s, const dims_t *d) {
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte;
  const wreal *p = s->p;
  wreal *cqw = s->cqw, *rt = s->rw_tend;
  for (int j = jts; j <= IMIN(jte, jde - 1); j++) {
    int k = kde;
    for (int i = its; i <= IMIN(ite, ide - 1); i++) {
      wreal cq1 = R(1.) / (R(1.) + A3(cqw, i, k - 1, j));
      wreal cq2 = A3(cqw, i, k - 1, j) * cq1;
      A3(rt, i, k, j) = A3(rt, i, k, j) + (R(1.) / A2(s->msfty, i, j)) * K_G *
                                              (cq1 * R(2.) * A1(s->rdnw, k - 1) * (-A3(p, i, k - 1, j)) - (A1(s->c1f, k) * A2(s->mu_2, i, j)) -
                                               cq2 * (A1(s->c1f, k) * A2(s->mub, i, j) + A1(s->c2f, k)));
    }
    for (k = 2; k <= kde - 1; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        wreal cq1 = R(1.) / (R(1.) + A3(cqw, i, k, j));
        wreal cq2 = A3(cqw, i, k, j) * cq1;
        A3(cqw, i, k, j) = cq1;
        A3(rt, i, k, j) = A3(rt, i, k, j) + (R(1.) / A2(s->msfty, i, j)) * K_G *
                                                (cq1 * A1(s->rdn, k) * (A3(p, i, k, j) - A3(p, i, k - 1, j)) - (A1(s->c1f, k) * A2(s->mu_2, i, j)) -
                                                 cq2 * (A1(s->c1f, k) * A2(s->mub, i, j) + A1(s->c2f, k)));
      }
  }
}

/* module_big_step_utilities_em.F:2503-2711; with w_damping = 0 only CFL diagnostics, which the oracle omits */
static void w_damp(ostate *s, const cfg_t *c, const dims_t *d) {
  if (c->w_damping != 1) return;
  const int ide = d->ide, jde = d->jde, kde = d->kde;
  wreal w_damp_on = c->zadvect_implicit > 0 ? (wreal)c->w_crit_cfl : K_W_BETA;
  for (int j = d->jts; j <= IMIN(d->jte, jde - 1); j++)
    for (int k = 2; k <= kde - 1; k++)
      for (int i = d->its; i <= IMIN(d->ite, ide - 1); i++) {
        wreal vert_cfl = (wreal)fabs((double)(A3(s->ww, i, k, j) / (A1(s->c1f, k) * A2(s->mut, i, j) + A1(s->c2f, k)) * A1(s->rdnw, k) * s->dt));
        if (vert_cfl > w_damp_on)
          A3(s->rw_tend, i, k, j) = A3(s->rw_tend, i, k, j) - sgn1(A3(s->w_2, i, k, j)) * K_W_ALPHA * (vert_cfl - (wreal)c->w_crit_cfl) *
                                                                  (A1(s->c1f, k) * A2(s->mut, i, j) + A1(s->c2f, k));
      }
}

/* module_big_step_utilities_em.F:3640-3850 */
static void coriolis(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  const int specified = c->specified || c->nested;
  const wreal *ru = s->ru, *rv = s->rv, *rw = s->rw, *f = s->f, *e = s->e, *sina = s->sina, *cosa = s->cosa;
  int ktf = IMIN(kte, kde - 1);
  int i_start = its, i_end = ite;
  if (c->open_xs || specified || c->nested) i_start = IMAX(ids + 1, its);
  if (c->open_xe || specified || c->nested) i_end = IMIN(ide - 1, ite);
  if (c->periodic_x) { i_start = its; i_end = ite; }
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        A3(s->ru_tend, i, k, j) = A3(s->ru_tend, i, k, j) +
            (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * R(0.5) * (A2(f, i, j) + A2(f, i - 1, j)) *
                R(0.25) * (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j)) -
            R(0.5) * (A2(e, i, j) + A2(e, i - 1, j)) * R(0.5) * (A2(cosa, i, j) + A2(cosa, i - 1, j)) *
                R(0.25) * (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  int j_start = jts, j_end = jte;
  if (c->open_ys || specified || c->nested || c->polar) j_start = IMAX(jds + 1, jts);
  if (c->open_ye || specified || c->nested || c->polar) j_end = IMIN(jde - 1, jte);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= IMIN(ide - 1, ite); i++)
        A3(s->rv_tend, i, k, j) = A3(s->rv_tend, i, k, j) -
            (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * R(0.5) * (A2(f, i, j) + A2(f, i, j - 1)) *
                R(0.25) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1)) +
            (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * R(0.5) * (A2(e, i, j) + A2(e, i, j - 1)) * R(0.5) * (A2(sina, i, j) + A2(sina, i, j - 1)) *
                R(0.25) * (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts + 1; k <= ktf; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++)
        A3(s->rw_tend, i, k, j) = A3(s->rw_tend, i, k, j) +
            A2(e, i, j) * (A2(cosa, i, j) * R(0.5) * (A1(s->fnm, k) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) +
                                                      A1(s->fnp, k) * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) -
                           (A2(s->msftx, i, j) / A2(s->msfty, i, j)) * A2(sina, i, j) * R(0.5) *
                               (A1(s->fnm, k) * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + A1(s->fnp, k) * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))));
}

/* module_big_step_utilities_em.F:4175-4469 (normal, non-polar code) */
static int curvature(ostate *s, const cfg_t *c, const dims_t *d) {
  const int ids = d->ids, ide = d->ide, jds = d->jds, jde = d->jde, kde = d->kde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  if (c->map_proj == 6 || c->polar) return WRFMI_E_BADOPT;
  const int specified = c->specified || c->nested;
  const wreal *ru = s->ru, *rv = s->rv, *rw = s->rw, *u = s->u_2, *v = s->v_2, *w = s->w_2;
  const wreal rdx = s->rdx, rdy = s->rdy;
  int ktf = IMIN(kte, kde - 1);
  /* vxgm(its-1:ite, kts:kte, jts-1:jte) */
  int vi0 = its - 1, vni = ite - its + 2, vnk = kte - kts + 1, vj0 = jts - 1;
  wreal *vxgm = (wreal *)calloc((size_t)vni * (size_t)vnk * (size_t)(jte - jts + 2), sizeof(wreal));
#define VX(i, k, j) vxgm[((size_t)((j) - vj0) * vnk + (size_t)((k) - kts)) * vni + (size_t)((i) - vi0)]
  int i_start = its - 1, i_end = ite, j_start = jts - 1, j_end = jte;
  if ((c->open_xs || specified || c->nested) && its == ids) i_start = its;
  if ((c->open_xe || specified || c->nested) && ite == ide) i_end = ite - 1;
  if ((c->open_ys || specified || c->nested || c->polar) && jts == jds) j_start = jts;
  if ((c->open_ye || specified || c->nested || c->polar) && jte == jde) j_end = jte - 1;
  if (c->periodic_x) { i_start = its - 1; i_end = ite; }
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        VX(i, k, j) = R(0.5) * (A3(u, i, k, j) + A3(u, i + 1, k, j)) * (A2(s->msfvx, i, j + 1) - A2(s->msfvx, i, j)) * rdy -
                      R(0.5) * (A3(v, i, k, j) + A3(v, i, k, j + 1)) * (A2(s->msfuy, i + 1, j) - A2(s->msfuy, i, j)) * rdx;
  if ((c->open_xs || (specified && !c->periodic_x) || c->nested) && its == ids)
    for (int j = jts; j <= jte - 1; j++)
      for (int k = kts; k <= ktf; k++) VX(its - 1, k, j) = VX(its, k, j);
  if ((c->open_xe || (specified && !c->periodic_x) || c->nested) && ite == ide)
    for (int j = jts; j <= jte - 1; j++)
      for (int k = kts; k <= ktf; k++) VX(ite, k, j) = VX(ite - 1, k, j);
  if ((c->open_ys || specified || c->nested || c->polar) && jts == jds)
    for (int k = kts; k <= ktf; k++)
      for (int i = its - 1; i <= ite; i++) VX(i, k, jts - 1) = VX(i, k, jts);
  if ((c->open_ye || specified || c->nested || c->polar) && jte == jde)
    for (int k = kts; k <= ktf; k++)
      for (int i = its - 1; i <= ite; i++) VX(i, k, jte) = VX(i, k, jte - 1);
  i_start = its;
  if (c->open_xs || specified || c->nested) i_start = IMAX(ids + 1, its);
  if (c->open_xe || specified || c->nested) i_end = IMIN(ide - 1, ite);
  if (c->periodic_x) { i_start = its; i_end = ite; }
  for (int j = jts; j <= IMIN(jde - 1, jte); j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = i_start; i <= i_end; i++)
        A3(s->ru_tend, i, k, j) = A3(s->ru_tend, i, k, j) +
            R(0.5) * (VX(i, k, j) + VX(i - 1, k, j)) * R(0.25) *
                (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j)) -
            A3(u, i, k, j) * K_RERADIUS * R(0.25) * (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  j_start = jts;
  if (c->open_ys || specified || c->nested || c->polar) j_start = IMAX(jds + 1, jts);
  if (c->open_ye || specified || c->nested || c->polar) j_end = IMIN(jde - 1, jte);
  for (int j = j_start; j <= j_end; j++)
    for (int k = kts; k <= ktf; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++)
        A3(s->rv_tend, i, k, j) = A3(s->rv_tend, i, k, j) -
            R(0.5) * (VX(i, k, j) + VX(i, k, j - 1)) * R(0.25) *
                (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1)) -
            (A2(s->msfvy, i, j) / A2(s->msfvx, i, j)) * A3(v, i, k, j) * K_RERADIUS * R(0.25) *
                (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = IMAX(2, kts); k <= ktf; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        wreal fm = A1(s->fnm, k), fp = A1(s->fnp, k);
        A3(s->rw_tend, i, k, j) = A3(s->rw_tend, i, k, j) +
            K_RERADIUS * (R(0.5) * (fm * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) + fp * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) *
                              R(0.5) * (fm * (A3(u, i, k, j) + A3(u, i + 1, k, j)) + fp * (A3(u, i, k - 1, j) + A3(u, i + 1, k - 1, j))) +
                          (A2(s->msftx, i, j) / A2(s->msfty, i, j)) * R(0.5) *
                              (fm * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + fp * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))) *
                              R(0.5) * (fm * (A3(v, i, k, j) + A3(v, i, k, j + 1)) + fp * (A3(v, i, k - 1, j) + A3(v, i, k - 1, j + 1))));
      }
#undef VX
  free(vxgm);
  return 0;
}

/* isotropic_km (dyn_em/module_diffusion_em.F:1718-1772, km_opt 1) with phy_bc's fill of the
 * halo: xkmh = khdif and xkhh = khdif/prandtl everywhere (prandtl = 1./3.0,
 * share/module_model_constants.F:86) */
static wreal *const_field(wreal v, const dims_t *d) {
  const size_t n = (size_t)NI_ * NK_ * (size_t)(d->jme - d->jms + 1);
  wreal *a = (wreal *)malloc(n * sizeof(wreal));
  for (size_t q = 0; q < n; q++) a[q] = v;
  return a;
}

/* dyn_em/module_em.F:799-877: diff_opt 1, the horizontal diffusion of u, v, w and theta
 * (theta relative to t_init) and, with no boundary-layer scheme, their vertical diffusion */
static void diff_opt1(ostate *s, const cfg_t *c, const dims_t *d) {
  const wreal khdif = (wreal)c->khdif, kvdif = (wreal)c->kvdif;
  wreal *xkmh = const_field(khdif, d), *xkhh = const_field(khdif / (R(1.) / R(3.0)), d);
#define HD(nm, fld, tnd, c1, c2) \
  o_horizontal_diffusion(nm, s->fld, s->tnd, s->mut, s->c1, s->c2, c, s->msfux, s->msfuy, s->msfvx, s->msfvx_inv, \
                         s->msfvy, s->msftx, s->msfty, khdif, xkmh, s->rdx, s->rdy, d)
  HD('u', u_2, ru_tendf, c1h, c2h);
  HD('v', v_2, rv_tendf, c1h, c2h);
  HD('w', w_2, rw_tendf, c1f, c2f);
#undef HD
  const wreal khdq = R(3.) * khdif;
  o_horizontal_diffusion_3dmp('m', s->t_2, s->t_tendf, s->mut, s->c1h, s->c2h, c, s->t_init, s->msfux, s->msfuy,
                              s->msfvx, s->msfvx_inv, s->msfvy, s->msftx, s->msfty, khdq, xkhh, s->rdx, s->rdy, d);
  if (c->bl_pbl_physics == 0) {
    o_vertical_diffusion_u(s->u_2, s->ru_tendf, c, s->u_base, s->c1h, s->c2h, s->alt, s->muu, s->rdn, s->rdnw, kvdif,
                           d);
    o_vertical_diffusion_v(s->v_2, s->rv_tendf, c, s->v_base, s->c1h, s->c2h, s->alt, s->muv, s->rdn, s->rdnw, kvdif,
                           d);
    if (c->non_hydrostatic)
      o_vertical_diffusion('w', s->w_2, s->rw_tendf, c, s->c1f, s->c2f, s->alt, s->mut, s->rdn, s->rdnw, kvdif, d);
    const wreal kvdq = R(3.) * kvdif;
    o_vertical_diffusion_3dmp(s->t_2, s->t_tendf, c, s->t_init, s->c1h, s->c2h, s->alt, s->mut, s->rdn, s->rdnw, kvdq,
                              d);
  }
  free(xkmh);
  free(xkhh);
}

/* dyn_em/module_em.F:190-955, physics off (rad_nudge = 0, IEVA off); at rk_step 1 the
 * forward-step terms of 797-950: diff_opt 1 with km_opt 1, sixth_order_diffusion
 * (diff_6th_opt 1, 2) and rk_rayleigh_damp (damp_opt 2) into the *_tendf arrays */
int o_rk_tendency(ostate *s, const cfg_t *c, const dims_t *d, int rk_step) {
  if (c->zadvect_implicit || c->momentum_adv_opt != 1 || c->diff_opt < 0 || c->diff_opt > 1 ||
      (c->diff_opt == 1 && c->km_opt != 1) || c->diff_6th_opt < 0 || c->diff_6th_opt > 2 || c->damp_opt == 1 ||
      c->pert_coriolis || d->kde + 1 >= WRFMI_ORACLE_MAXK)
    return WRFMI_E_BADOPT;
  zero3(s->ru_tend, d); zero3(s->rv_tend, d); zero3(s->rw_tend, d); zero3(s->t_tend, d); zero3(s->ph_tend, d);
  zero3(s->u_save, d); zero3(s->v_save, d); zero3(s->w_save, d); zero3(s->ph_save, d); zero3(s->t_save, d);
  zero2(s->mu_tend, d); zero2(s->mu_save, d);
  int ts = c->time_step, rc;
  /* wwE == ww wherever the advection reads it (WW_SPLIT copy, module_ieva_em.F:294-302) */
  rc = o_advect_u(s->u_2, s->u_2, s->ru_tend, s->ru, s->rv, s->ww, s->c1h, s->c2h, s->mut, ts, c, s->msfux, s->msfuy,
                  s->msfvx, s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp, s->rdx, s->rdy, s->rdnw, d);
  if (rc) return rc;
  rc = o_advect_v(s->v_2, s->v_2, s->rv_tend, s->ru, s->rv, s->ww, s->c1h, s->c2h, s->mut, ts, c, s->msfux, s->msfuy,
                  s->msfvx, s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp, s->rdx, s->rdy, s->rdnw, d);
  if (rc) return rc;
  if (c->non_hydrostatic) {
    rc = o_advect_w(s->w_2, s->w_2, s->rw_tend, s->ru, s->rv, s->ww, s->c1h, s->c2h, s->mut, ts, c, s->msfux, s->msfuy,
                    s->msfvx, s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp, s->rdx, s->rdy, s->rdn, d);
    if (rc) return rc;
  }
  rc = o_advect_scalar(s->t_2, s->t_2, s->t_tend, s->ru, s->rv, s->ww, s->c1h, s->c2h, s->mut, ts, c, s->msfux,
                       s->msfuy, s->msfvx, s->msfvy, s->msftx, s->msfty, s->fnm, s->fnp, s->rdx, s->rdy, s->rdnw, d);
  if (rc) return rc;
  rhs_ph(s, c, d);
  horizontal_pressure_gradient(s, c, d);
  if (c->non_hydrostatic) pg_buoy_w(s, d);
  w_damp(s, c, d);
  coriolis(s, c, d);
  rc = curvature(s, c, d);
  if (rc || rk_step != 1) return rc;
  if (c->diff_opt == 1) diff_opt1(s, c, d);
  if (c->diff_6th_opt != 0) {
    const wreal f6 = (wreal)c->diff_6th_factor;
#define SIX(nm, fld, tnd, c1, c2) \
    o_sixth_order_diffusion(nm, s->fld, s->tnd, s->mut, s->dt, c, s->c1, s->c2, c->diff_6th_opt, f6, s->phb, s->ph_2, \
                            s->rdx, s->rdy, s->msftx, s->msfty, s->msfux, s->msfuy, s->msfvx, s->msfvy, d)
    SIX('u', u_2, ru_tendf, c1h, c2h);
    SIX('v', v_2, rv_tendf, c1h, c2h);
    if (c->non_hydrostatic) SIX('w', w_2, rw_tendf, c1f, c2f);
    SIX('m', t_2, t_tendf, c1h, c2h);
#undef SIX
  }
  if (c->damp_opt == 2) o_rk_rayleigh_damp(s, c, d);
  return 0;
}

/* dyn_em/module_em.F:959-1092 (c1 = c1h, c2 = c2h) */
void o_rk_addtend_dry(ostate *s, const dims_t *d, int rk_step) {
  const int ide = d->ide, jde = d->jde;
  const int its = d->its, ite = d->ite, jts = d->jts, jte = d->jte, kts = d->kts, kte = d->kte;
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= kte - 1; k++)
      for (int i = its; i <= ite; i++) {
        if (rk_step == 1) A3(s->ru_tendf, i, k, j) = A3(s->ru_tendf, i, k, j) + A3(s->u_save, i, k, j) * A2(s->msfuy, i, j);
        A3(s->ru_tend, i, k, j) = A3(s->ru_tend, i, k, j) + A3(s->ru_tendf, i, k, j) / A2(s->msfuy, i, j);
      }
  for (int j = jts; j <= jte; j++)
    for (int k = kts; k <= kte - 1; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        if (rk_step == 1) A3(s->rv_tendf, i, k, j) = A3(s->rv_tendf, i, k, j) + A3(s->v_save, i, k, j) * A2(s->msfvx, i, j);
        A3(s->rv_tend, i, k, j) = A3(s->rv_tend, i, k, j) + A3(s->rv_tendf, i, k, j) * A2(s->msfvx_inv, i, j);
      }
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= kte; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        if (rk_step == 1) A3(s->rw_tendf, i, k, j) = A3(s->rw_tendf, i, k, j) + A3(s->w_save, i, k, j) * A2(s->msfty, i, j);
        A3(s->rw_tend, i, k, j) = A3(s->rw_tend, i, k, j) + A3(s->rw_tendf, i, k, j) / A2(s->msfty, i, j);
        if (rk_step == 1) A3(s->ph_tendf, i, k, j) = A3(s->ph_tendf, i, k, j) + A3(s->ph_save, i, k, j);
        A3(s->ph_tend, i, k, j) = A3(s->ph_tend, i, k, j) + A3(s->ph_tendf, i, k, j) / A2(s->msfty, i, j);
      }
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int k = kts; k <= kte - 1; k++)
      for (int i = its; i <= IMIN(ite, ide - 1); i++) {
        if (rk_step == 1) A3(s->t_tendf, i, k, j) = A3(s->t_tendf, i, k, j) + A3(s->t_save, i, k, j);
        A3(s->t_tend, i, k, j) = A3(s->t_tend, i, k, j) + A3(s->t_tendf, i, k, j) / A2(s->msfty, i, j) +
                                 (A1(s->c1h, k) * A2(s->mut, i, j) + A1(s->c2h, k)) * A3(s->h_diabatic, i, k, j) / A2(s->msfty, i, j);
      }
  for (int j = jts; j <= IMIN(jte, jde - 1); j++)
    for (int i = its; i <= IMIN(ite, ide - 1); i++) A2(s->mu_tend, i, j) = A2(s->mu_tend, i, j) + A2(s->mu_tendf, i, j);
}
