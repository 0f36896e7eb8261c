/*
 * wrfmi_fields.h -- the field list of one WRF-ARW dynamics domain, as X-macros.
 *
 * Every array below is a Fortran explicit-shape array of the reference:
 *   3-D  REAL, DIMENSION(ims:ime, kms:kme, jms:jme)   (i fastest, then k, then j)
 *   2-D  REAL, DIMENSION(ims:ime, jms:jme)
 *   1-D  REAL, DIMENSION(kms:kme)
 * Names are the reference's grid%<name> fields (Registry/Registry.EM_COMMON:158-463)
 * or the solve_em locals of the same name (dyn_em/solve_em.F, "i1" Registry arrays),
 * so a host that owns a WRF domain hands its own arrays over unchanged.
 *
 * The list is consumed three ways: include/wrfmi.h builds the C-ABI state struct,
 * the CPU oracle builds its own struct with the same member order, and the Python
 * host parses this file to build its ctypes mirror.  Keep one name per X().
 */
#ifndef WRFMI_FIELDS_H
#define WRFMI_FIELDS_H

/* 3-D fields (mass, u-, v- or w-staggered; staggering is implied by the name) */
#define WRFMI_FIELDS_3D(X)                                                      \
  X(u_1) X(u_2) X(v_1) X(v_2) X(w_1) X(w_2) X(t_1) X(t_2) X(ph_1) X(ph_2)      \
  X(phb) X(pb) X(alb) X(p) X(al) X(alt) X(php)                                  \
  X(ru) X(rv) X(rw) X(ww) X(ru_m) X(rv_m) X(ww_m)                               \
  X(u_save) X(v_save) X(w_save) X(t_save) X(ph_save)                            \
  X(ru_tend) X(rv_tend) X(rw_tend) X(ph_tend) X(t_tend)                         \
  X(ru_tendf) X(rv_tendf) X(rw_tendf) X(ph_tendf) X(t_tendf)                    \
  X(h_diabatic)                                                                 \
  X(cqu) X(cqv) X(cqw) X(c2a) X(pm1) X(ww1) X(t_2save) X(advect_tend)            \
  X(t_init)

/* 2-D fields */
#define WRFMI_FIELDS_2D(X)                                                      \
  X(mu_1) X(mu_2) X(mub) X(mut) X(muu) X(muv) X(muts) X(muus) X(muvs)           \
  X(mudf) X(muave) X(mu_save) X(mu_tend) X(mu_tendf) X(ht)                      \
  X(msfux) X(msfuy) X(msfvx) X(msfvx_inv) X(msfvy) X(msftx) X(msfty)            \
  X(f) X(e) X(sina) X(cosa) X(xlat)

/* 1-D vertical arrays: level coefficients, the base-state profiles of the upper-level
 * Rayleigh damper (u, v, theta perturbation and height of a reference column; u_base and
 * v_base are also the reference of diff_opt 1's vertical diffusion) and the QV profile
 * that diff_opt 1 diffuses QV relative to */
#define WRFMI_FIELDS_1D(X)                                                      \
  X(znu) X(znw) X(dnw) X(rdnw) X(dn) X(rdn) X(fnm) X(fnp)                       \
  X(c1h) X(c2h) X(c1f) X(c2f) X(c3h) X(c4h) X(c3f) X(c4f)                       \
  X(u_base) X(v_base) X(t_base) X(z_base) X(qv_base)

/* REAL scalars of the domain (grid%<name>) */
#define WRFMI_SCALARS_R(X)                                                      \
  X(dt) X(dx) X(dy) X(rdx) X(rdy) X(cf1) X(cf2) X(cf3) X(cfn) X(cfn1) X(p_top)

/* maximum number of advected species (moist first, then passive scalars) */
#define WRFMI_MAX_SPECIES 16

#endif
