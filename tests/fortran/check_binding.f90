! Links a Fortran host against libwrfmi.so through wrf_amd/fortran/wrfmi_bindc.f90 and checks
! the struct layouts and the option check of the step entry point (no GPU needed).
program check_binding
  use iso_c_binding
  use wrfmi_bindc
  implicit none
  interface
    integer(c_size_t) function wrfmi_state_sizeof() bind(C, name="wrfmi_state_sizeof")
      import :: c_size_t
    end function
  end interface
  type(wrfmi_state) :: st
  type(wrfmi_config) :: cfg
  type(wrfmi_dims) :: d
  integer(c_int) :: rc
  character(kind=c_char), pointer :: msg(:)
  integer :: n
  if (c_sizeof(st) /= wrfmi_state_sizeof()) then
    print *, "FAIL state size", c_sizeof(st), wrfmi_state_sizeof()
    stop 1
  end if
  cfg%rk_ord = 2
  rc = wrfmi_solve_em_step(st, cfg, d, c_null_ptr, 1_c_int)
  call c_f_pointer(wrfmi_last_error(), msg, [200])
  n = 1
  do while (msg(n) /= c_null_char .and. n < 200)
    n = n + 1
  end do
  if (rc /= 1) then
    print *, "FAIL expected WRFMI_E_BADOPT, got", rc
    stop 1
  end if
  print *, "OK ", c_sizeof(st), " bytes; ", msg(1:n-1)
end program check_binding
