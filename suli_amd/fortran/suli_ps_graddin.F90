! Drop-in graddin(), pressh(), posdin() and exp_press() for SuLi on the
! MI355X library.
!
! Replaces SuLi-IC/9_grad.f90 (graddin, :10-663; pressh, :667-702),
! SuLi-IC/10_pos.f90 (posdin, :1-99) and SuLi-IC/11_diversos.f90
! (exp_press, :455-495: velpre's prd0, px, py, pz) in a SuLi build (the
! reference's exp_press is then taken out of 11_diversos.f90): same names, no
! arguments, same module variables read and written (velpre: u, v, w, rho,
! prd1, prd; disc: nx..dt, t_press, obst_t, ibm_t; cond: ccx0, ccy0; obst:
! obs_ls*; ls_param: rho_f1, hs; param: gz).  Compile with -DSULI_ICDZ for SuLi-ICdz, whose disc module holds
! the dz(0:nz+1) and dzz(0:nz+1) arrays.
!
! Run-time options (environment, read once):
!   SULI_PS_PRECISION  fp64 (default) | mixed
!   SULI_PS_TOL        relative tolerance; unset = the reference stop rule
!                      r.r <= 1e-4/(nx*ny*nz) (SuLi-IC/9_grad.f90:393)
!   SULI_PS_MAXITER    iteration cap (default 10000, as the reference)
module suli_ps_dropin_state
   use iso_c_binding
   use suli_ps_iface
   implicit none
   type(c_ptr), save :: suli_ps_ctx = c_null_ptr
contains
   subroutine suli_ps_dropin_init()
      type(suli_ps_opts) :: o
      character(len=64) :: val
      integer :: n, ios
      if (c_associated(suli_ps_ctx)) return
      call get_environment_variable("SULI_PS_PRECISION", val, length=n)
      if (n > 0 .and. trim(val) == "mixed") o%precision = SULI_PS_MIXED
      call get_environment_variable("SULI_PS_TOL", val, length=n)
      if (n > 0) then
         read (val, *, iostat=ios) o%tol
         if (ios == 0) o%stop_rule = SULI_PS_STOP_RELATIVE
      end if
      call get_environment_variable("SULI_PS_MAXITER", val, length=n)
      if (n > 0) then
         read (val, *, iostat=ios) o%max_iter
         if (ios /= 0) o%max_iter = 10000
      end if
      o%verbose = 0
      ! SuLi's loop calls posdin() right after graddin() on the same module
      ! arrays (1_principal.f90:102-103): posdin takes u, v, w, rho and prd1
      ! from the device copies graddin left instead of uploading them again
      o%reuse_staged = 1
      suli_ps_ctx = suli_ps_create(o)
      if (.not. c_associated(suli_ps_ctx)) then
         write (*, *) "suli_ps: no HIP device / context creation failed"
         error stop 1
      end if
   end subroutine suli_ps_dropin_init

   subroutine suli_ps_dropin_grid(g)
      use disc
      use cond
      use ls_param, only: rho_f1
      type(suli_ps_grid), intent(out) :: g
      g%nx = nx; g%ny = ny; g%nz = nz
      g%ccx0 = ccx0; g%ccy0 = ccy0
      g%t_press = t_press
      g%obst_t = obst_t; g%ibm_t = ibm_t
      g%dx = dx; g%dy = dy
#ifdef SULI_ICDZ
      g%dz = dz(1)
#else
      g%dz = dz
#endif
      g%dt = dt
      g%rho_f1 = rho_f1
   end subroutine suli_ps_dropin_grid

   subroutine suli_ps_dropin_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc < 0) then
         write (*, *) "suli_ps ", where, ": ", trim(suli_ps_error_message(suli_ps_ctx))
         error stop 1
      end if
   end subroutine suli_ps_dropin_check
end module suli_ps_dropin_state


SUBROUTINE graddin()
   use iso_c_binding
   use velpre
   use obst
   use suli_ps_iface
   use suli_ps_dropin_state
   implicit none
   type(suli_ps_grid) :: g
   type(suli_ps_stats) :: st
   integer(c_int) :: rc

   call suli_ps_dropin_init()
   call suli_ps_dropin_grid(g)
   if (obst_t /= 0 .and. ibm_t == 2) then
#ifdef SULI_ICDZ
      rc = suli_ps_graddin(suli_ps_ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, &
                           prd1=prd1, st=st)
#else
      rc = suli_ps_graddin(suli_ps_ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, &
                           prd1=prd1, st=st)
#endif
   else
#ifdef SULI_ICDZ
      rc = suli_ps_graddin(suli_ps_ctx, g, u, v, w, rho, dz=dz, dzz=dzz, prd1=prd1, st=st)
#else
      rc = suli_ps_graddin(suli_ps_ctx, g, u, v, w, rho, prd1=prd1, st=st)
#endif
   end if
   call suli_ps_dropin_check(rc, "graddin")
   ! the reference's warning when CG hits its cap (SuLi-IC/9_grad.f90:590)
   if (rc == 1) write (*, *) "pulou pressão; ", "erro =", st%rr
END SUBROUTINE graddin


SUBROUTINE posdin()
   use iso_c_binding
   use velpre
   use ls_param, only: hs
   use suli_ps_iface
   use suli_ps_dropin_state
   implicit none
   type(suli_ps_grid) :: g
   integer(c_int) :: rc

   call suli_ps_dropin_init()
   call suli_ps_dropin_grid(g)
   if (t_press /= 1) then
      ! the hydrostatic branch (SuLi-IC/10_pos.f90:47-94) reads hs
#ifdef SULI_ICDZ
      rc = suli_ps_posdin_hydro(suli_ps_ctx, g, u, v, w, rho, hs, prd1, dz, prd)
#else
      rc = suli_ps_posdin_hydro(suli_ps_ctx, g, u, v, w, rho, hs, prd1, prd=prd)
#endif
   else
#ifdef SULI_ICDZ
      rc = suli_ps_posdin(suli_ps_ctx, g, u, v, w, rho, prd1, dz, dzz, prd)
#else
      rc = suli_ps_posdin(suli_ps_ctx, g, u, v, w, rho, prd1, prd=prd)
#endif
   end if
   call suli_ps_dropin_check(rc, "posdin")
END SUBROUTINE posdin


SUBROUTINE pressh()
   use iso_c_binding
   use velpre
   use param, only: gz
   use suli_ps_iface
   use suli_ps_dropin_state
   implicit none
   type(suli_ps_grid) :: g
   integer(c_int) :: rc

   call suli_ps_dropin_init()
   call suli_ps_dropin_grid(g)
#ifdef SULI_ICDZ
   rc = suli_ps_pressh(suli_ps_ctx, g, rho, dz, gz, prd1)
#else
   rc = suli_ps_pressh(suli_ps_ctx, g, rho, gz=gz, prd1=prd1)
#endif
   call suli_ps_dropin_check(rc, "pressh")
END SUBROUTINE pressh


SUBROUTINE exp_press()
   use iso_c_binding
   use velpre
   use suli_ps_iface
   use suli_ps_dropin_state
   implicit none
   type(suli_ps_grid) :: g
   integer(c_int) :: rc

   call suli_ps_dropin_init()
   call suli_ps_dropin_grid(g)
#ifdef SULI_ICDZ
   rc = suli_ps_exp_press(suli_ps_ctx, g, rho, prd1, prd0, dz, dzz, px, py, pz)
#else
   rc = suli_ps_exp_press(suli_ps_ctx, g, rho, prd1, prd0, px=px, py=py, pz=pz)
#endif
   call suli_ps_dropin_check(rc, "exp_press")
END SUBROUTINE exp_press
