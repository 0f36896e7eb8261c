! A solve_em-shaped Fortran host of the MI355X dynamics step: one RK3 time step of
! dyn_em/solve_em.F (578-4740) for the implemented options, written as a maintainer would
! change solve_em to call the GPU library.
!
! The acoustic loop and the scalar update call the entry points that keep the reference's
! argument lists, each at its own call site and with the grid fields solve_em passes there:
! advance_uv (solve_em.F:1279), advance_mu_t (1392), advance_w (1499), sumflux (1566),
! calc_p_rho (1623) and rk_update_scalar (2309), followed by the boundary fills of
! solve_em.F:1640-1719.  The routines whose argument lists span the whole grid (rk_step_prep,
! rk_tendency, small_step_prep with calc_p_rho and calc_coef_w, calc_mu_uv_1 with
! small_step_finish, rk_scalar_tend, calc_p_rho_phi and the boundary-condition groups) go
! through the state-struct entry points.
!
! The grid is the wrfmi_state struct: c_ptr members that this host turns into Fortran
! arrays (contiguous pointers, so that an array argument passes its own address) and never
! dereferences; they may be device arrays or host arrays registered with the library.
module wrfmi_solve_em_host
  use iso_c_binding
  use wrfmi_bindc
  implicit none
  private
  public :: wrfmi_fortran_solve_em_step

contains

  ! number of acoustic steps per time step (solve_em.F:446-463)
  integer function sound_steps(cfg, st)
    type(wrfmi_config), intent(in) :: cfg
    type(wrfmi_state), intent(in) :: st
    real(c_float) :: spacing
    if (cfg%time_step_sound /= 0) then
      sound_steps = cfg%time_step_sound
    else
      spacing = min(st%dx, st%dy)
      sound_steps = 2 * (int(300. * st%dt / spacing - 0.01) + 1)
      sound_steps = max(sound_steps, 4)
    end if
  end function sound_steps

  integer(c_int) function wrfmi_fortran_solve_em_step(st, cfg, dm) bind(C, name="wrfmi_fortran_solve_em_step")
    type(wrfmi_state), intent(inout) :: st
    type(wrfmi_config), intent(in) :: cfg
    type(wrfmi_dims), intent(in) :: dm
    real(c_float), pointer, contiguous :: u_2(:), v_2(:), w_2(:), t_2(:), ph_2(:), p(:), pb(:), al(:), alt(:), &
        php(:), ru_tend(:), rv_tend(:), rw_tend(:), ph_tend(:), t_tend(:), cqu(:), cqv(:), cqw(:), c2a(:), &
        pm1(:), ww(:), ww1(:), ru_m(:), rv_m(:), ww_m(:), u_save(:), v_save(:), w_save(:), t_save(:), &
        ph_save(:), t_2save(:), phb(:), alb(:), advect_tend(:), a(:), alpha(:), gamma(:)
    real(c_float), pointer, contiguous :: mu_1(:), mu_2(:), mub(:), mut(:), muu(:), muv(:), muts(:), mudf(:), &
        muave(:), mu_tend(:), ht(:), msfux(:), msfuy(:), msfvx(:), msfvx_inv(:), msfvy(:), msftx(:), msfty(:)
    real(c_float), pointer, contiguous :: znu(:), dnw(:), rdnw(:), rdn(:), fnm(:), fnp(:), c1h(:), c2h(:), &
        c1f(:), c2f(:), c3h(:), c4h(:), c3f(:), c4f(:)
    real(c_float), pointer, contiguous :: q(:), q_old(:), q_tend(:)
    integer(c_int) :: ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte
    integer(c_int) :: rk_step, iteration, number_of_small_timesteps, im, spec_zone, nh, top_lid, one, tenddec
    integer :: n_sound, n3, n2, n1
    real(c_float) :: dt_rk, dts_rk, t0
    type(c_ptr), parameter :: stream = c_null_ptr

    ids = dm%ids; ide = dm%ide; jds = dm%jds; jde = dm%jde; kds = dm%kds; kde = dm%kde
    ims = dm%ims; ime = dm%ime; jms = dm%jms; jme = dm%jme; kms = dm%kms; kme = dm%kme
    its = dm%its; ite = dm%ite; jts = dm%jts; jte = dm%jte; kts = dm%kts; kte = dm%kte
    n3 = (ime - ims + 1) * (kme - kms + 1) * (jme - jms + 1)
    n2 = (ime - ims + 1) * (jme - jms + 1)
    n1 = kme - kms + 1
    call c_f_pointer(st%u_2, u_2, [n3]);         call c_f_pointer(st%v_2, v_2, [n3])
    call c_f_pointer(st%w_2, w_2, [n3]);         call c_f_pointer(st%t_2, t_2, [n3])
    call c_f_pointer(st%ph_2, ph_2, [n3]);       call c_f_pointer(st%p, p, [n3])
    call c_f_pointer(st%pb, pb, [n3]);           call c_f_pointer(st%al, al, [n3])
    call c_f_pointer(st%alt, alt, [n3]);         call c_f_pointer(st%php, php, [n3])
    call c_f_pointer(st%ru_tend, ru_tend, [n3]); call c_f_pointer(st%rv_tend, rv_tend, [n3])
    call c_f_pointer(st%rw_tend, rw_tend, [n3]); call c_f_pointer(st%ph_tend, ph_tend, [n3])
    call c_f_pointer(st%t_tend, t_tend, [n3])
    call c_f_pointer(st%cqu, cqu, [n3]);         call c_f_pointer(st%cqv, cqv, [n3])
    call c_f_pointer(st%cqw, cqw, [n3]);         call c_f_pointer(st%c2a, c2a, [n3])
    call c_f_pointer(st%pm1, pm1, [n3]);         call c_f_pointer(st%ww, ww, [n3])
    call c_f_pointer(st%ww1, ww1, [n3]);         call c_f_pointer(st%ru_m, ru_m, [n3])
    call c_f_pointer(st%rv_m, rv_m, [n3]);       call c_f_pointer(st%ww_m, ww_m, [n3])
    call c_f_pointer(st%u_save, u_save, [n3]);   call c_f_pointer(st%v_save, v_save, [n3])
    call c_f_pointer(st%w_save, w_save, [n3]);   call c_f_pointer(st%t_save, t_save, [n3])
    call c_f_pointer(st%ph_save, ph_save, [n3]); call c_f_pointer(st%t_2save, t_2save, [n3])
    call c_f_pointer(st%phb, phb, [n3]);         call c_f_pointer(st%alb, alb, [n3])
    call c_f_pointer(st%advect_tend, advect_tend, [n3])
    call c_f_pointer(st%a, a, [n3]);             call c_f_pointer(st%alpha, alpha, [n3])
    call c_f_pointer(st%gamma, gamma, [n3])
    call c_f_pointer(st%mu_1, mu_1, [n2]);       call c_f_pointer(st%mu_2, mu_2, [n2])
    call c_f_pointer(st%mub, mub, [n2]);         call c_f_pointer(st%mut, mut, [n2])
    call c_f_pointer(st%muu, muu, [n2]);         call c_f_pointer(st%muv, muv, [n2])
    call c_f_pointer(st%muts, muts, [n2]);       call c_f_pointer(st%mudf, mudf, [n2])
    call c_f_pointer(st%muave, muave, [n2]);     call c_f_pointer(st%mu_tend, mu_tend, [n2])
    call c_f_pointer(st%ht, ht, [n2])
    call c_f_pointer(st%msfux, msfux, [n2]);     call c_f_pointer(st%msfuy, msfuy, [n2])
    call c_f_pointer(st%msfvx, msfvx, [n2]);     call c_f_pointer(st%msfvx_inv, msfvx_inv, [n2])
    call c_f_pointer(st%msfvy, msfvy, [n2]);     call c_f_pointer(st%msftx, msftx, [n2])
    call c_f_pointer(st%msfty, msfty, [n2])
    call c_f_pointer(st%znu, znu, [n1]);         call c_f_pointer(st%dnw, dnw, [n1])
    call c_f_pointer(st%rdnw, rdnw, [n1]);       call c_f_pointer(st%rdn, rdn, [n1])
    call c_f_pointer(st%fnm, fnm, [n1]);         call c_f_pointer(st%fnp, fnp, [n1])
    call c_f_pointer(st%c1h, c1h, [n1]);         call c_f_pointer(st%c2h, c2h, [n1])
    call c_f_pointer(st%c1f, c1f, [n1]);         call c_f_pointer(st%c2f, c2f, [n1])
    call c_f_pointer(st%c3h, c3h, [n1]);         call c_f_pointer(st%c4h, c4h, [n1])
    call c_f_pointer(st%c3f, c3f, [n1]);         call c_f_pointer(st%c4f, c4f, [n1])

    spec_zone = 1
    nh = cfg%non_hydrostatic
    top_lid = cfg%top_lid
    one = 1
    tenddec = 0
    t0 = 300.
    n_sound = sound_steps(cfg, st)

    wrfmi_fortran_solve_em_step = 0
    rk_loop: do rk_step = 1, 3
      ! substep lengths, solve_em.F:601-615
      if (rk_step == 1) then
        dt_rk = st%dt / 3.
        dts_rk = dt_rk
        number_of_small_timesteps = 1
      else if (rk_step == 2) then
        dt_rk = 0.5 * st%dt
        dts_rk = st%dt / real(n_sound)
        number_of_small_timesteps = n_sound / 2
      else
        dt_rk = st%dt
        dts_rk = st%dt / real(n_sound)
        number_of_small_timesteps = n_sound
      end if

      if (chk(wrfmi_rk_step_prep(st, cfg, dm, rk_step, stream))) exit rk_loop
      if (chk(wrfmi_rk_tendency(st, cfg, dm, rk_step, stream))) exit rk_loop
      if (chk(wrfmi_small_step_prep(st, cfg, dm, rk_step, dts_rk, stream))) exit rk_loop

      do iteration = 1, number_of_small_timesteps
        if (chk(wrfmi_advance_uv(u_2, ru_tend, v_2, rv_tend, p, pb, ph_2, php, alt, al, mu_2, muu, cqu, muv, &
            cqv, mudf, c1h, c2h, c1f, c2f, c3h, c4h, c3f, c4f, msfux, msfuy, msfvx, msfvx_inv, msfvy, &
            st%rdx, st%rdy, dts_rk, st%cf1, st%cf2, st%cf3, fnm, fnp, cfg%emdiv, rdnw, cfg, spec_zone, nh, &
            top_lid, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) &
            exit rk_loop
        if (chk(wrfmi_advance_mu_t(ww, ww1, u_2, u_save, v_2, v_save, mu_2, mut, muave, muts, muu, muv, mudf, &
            c1h, c2h, c1f, c2f, c3h, c4h, c3f, c4f, ru_m, rv_m, ww_m, t_2, t_save, t_2save, t_tend, mu_tend, &
            st%rdx, st%rdy, dts_rk, cfg%epssm, dnw, fnm, fnp, rdnw, msfux, msfuy, msfvx, msfvx_inv, msfvy, &
            msftx, msfty, iteration, cfg, &
            ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) exit rk_loop
        if (chk(wrfmi_advance_w(w_2, rw_tend, ww, w_save, u_2, v_2, mu_2, mut, muave, muts, c1h, c2h, c1f, &
            c2f, c3h, c4h, c3f, c4f, t_2save, t_2, t_save, ph_2, ph_save, phb, ph_tend, ht, c2a, cqw, alt, &
            alb, a, alpha, gamma, st%rdx, st%rdy, dts_rk, t0, cfg%epssm, dnw, fnm, fnp, rdnw, rdn, st%cf1, &
            st%cf2, st%cf3, msftx, msfty, cfg, top_lid, &
            ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) exit rk_loop
        if (chk(wrfmi_sumflux(u_2, v_2, ww, u_save, v_save, ww1, muu, muv, c1h, c2h, c1f, c2f, c3h, c4h, c3f, &
            c4f, ru_m, rv_m, ww_m, cfg%epssm, msfux, msfuy, msfvx, msfvx_inv, msfvy, iteration, &
            number_of_small_timesteps, &
            ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) exit rk_loop
        if (chk(wrfmi_calc_p_rho(al, p, ph_2, alt, t_2, t_save, c2a, pm1, mu_2, muts, c1h, c2h, c1f, c2f, c3h, &
            c4h, c3f, c4f, znu, t0, rdnw, dnw, cfg%smdiv, nh, iteration, &
            ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) exit rk_loop
        ! the boundary fills after the acoustic substep (solve_em.F:1640-1719)
        if (chk(wrfmi_set_physical_bc3d(st%ph_2, 'w', cfg, dm, stream))) exit rk_loop
        if (chk(wrfmi_set_physical_bc3d(st%al, 'p', cfg, dm, stream))) exit rk_loop
        if (chk(wrfmi_set_physical_bc3d(st%p, 'p', cfg, dm, stream))) exit rk_loop
        if (chk(wrfmi_set_physical_bc2d(st%muts, 't', cfg, dm, stream))) exit rk_loop
        if (chk(wrfmi_set_physical_bc2d(st%mu_2, 't', cfg, dm, stream))) exit rk_loop
        if (chk(wrfmi_set_physical_bc2d(st%mudf, 't', cfg, dm, stream))) exit rk_loop
      end do

      if (chk(wrfmi_small_step_finish(st, cfg, dm, rk_step, number_of_small_timesteps, dts_rk, stream))) exit rk_loop

      ! positive-definite transport (moist_adv_opt / scalar_adv_opt 1): before the last
      ! substep's transport, each such species takes its physics tendency into its time-n
      ! values with mu_1 as both masses, and those get their halo (solve_em.F:1850-1975)
      if (rk_step == 3) then
        do im = 1, st%n_species
          if (.not. pd_species(im)) cycle
          call c_f_pointer(st%q_old(im), q_old, [n3])
          call c_f_pointer(st%q_tend(im), q_tend, [n3])
          if (chk(wrfmi_rk_update_scalar_pd(one, one, q_old, q_tend, c1h, c2h, mu_1, mu_1, mub, rk_step, dt_rk, &
              spec_zone, cfg, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, &
              kts, kte))) exit rk_loop
          if (chk(wrfmi_set_physical_bc3d(st%q_old(im), 'p', cfg, dm, stream))) exit rk_loop
        end do
      end if

      ! moist and passive species: rk_scalar_tend (advect_scalar_pd on the last substep for a
      ! positive-definite species), then rk_update_scalar for one species
      ! (solve_em.F:2215-2309, 2781-2856)
      do im = 1, st%n_species
        if (chk(wrfmi_rk_scalar_tend(st, cfg, dm, im - 1, rk_step, dt_rk, stream))) exit rk_loop
        call c_f_pointer(st%q(im), q, [n3])
        call c_f_pointer(st%q_old(im), q_old, [n3])
        call c_f_pointer(st%q_tend(im), q_tend, [n3])
        if (chk(wrfmi_rk_update_scalar(one, one, q_old, q, q_tend, c_null_ptr, c_null_ptr, advect_tend, &
            advect_tend, advect_tend, msftx, msfty, c1h, c2h, mu_1, mu_2, mub, rk_step, dt_rk, spec_zone, cfg, &
            tenddec, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))) &
            exit rk_loop
      end do

      if (chk(wrfmi_calc_p_rho_phi(st, cfg, dm, stream))) exit rk_loop
      if (rk_step < 3) then
        if (chk(wrfmi_phys_bc_dry_2(st, cfg, dm, 0, stream))) exit rk_loop
      end if
    end do rk_loop
    if (wrfmi_fortran_solve_em_step /= 0) return

    ! after the RK loop: calc_p_rho_phi again (solve_em.F:4195), set_phys_bc_dry_2 and the
    ! species fills (4461-4540), and the surface w (4740)
    if (chk(wrfmi_calc_p_rho_phi(st, cfg, dm, stream))) return
    if (chk(wrfmi_phys_bc_dry_2(st, cfg, dm, 1, stream))) return
    if (chk(wrfmi_set_w_surface(st, cfg, dm, stream))) return

  contains

    ! species im (1-based; moist species first) is advected in positive-definite form
    logical function pd_species(im)
      integer(c_int), intent(in) :: im
      if (im <= st%n_moist) then
        pd_species = cfg%moist_adv_opt == 1
      else
        pd_species = cfg%scalar_adv_opt == 1
      end if
    end function pd_species

    ! true (and the status kept as the function result) when a call failed: the reference's
    ! host would call wrf_error_fatal with wrfmi_last_error() here
    logical function chk(rc)
      integer(c_int), intent(in) :: rc
      chk = rc /= 0
      if (chk) wrfmi_fortran_solve_em_step = rc
    end function chk

  end function wrfmi_fortran_solve_em_step

end module wrfmi_solve_em_host
