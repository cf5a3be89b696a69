! Drop-in graddin() for SuLi-IT on the MI355X library (posdin() is in
! suli_ps_posdin_it.F90).
!
! Replaces SuLi-IT/9_grad.f90 (graddin, :10-263) and SuLi-IT/10_pos.f90
! (posdin, :73-132) in a SuLi-IT build: same names, no arguments, the same
! module variables (disc: nx..dt, dz; cond: ccx0, ccy0; velpre: u, v, w,
! prd1, prd, eta0, eta1, deltah; dzs: dzx, dzy, dzz; parametros: tetah, gz).
! gradeta() and the rest of 9_grad.f90's hydrostatic part stay the
! reference's.  Run-time options as suli_ps_graddin.F90 (SULI_PS_PRECISION,
! SULI_PS_TOL, SULI_PS_MAXITER).
module suli_ps_it_state
   use iso_c_binding
   use suli_ps_iface
   implicit none
   type(c_ptr), save :: suli_ps_it_ctx = c_null_ptr
contains
   subroutine suli_ps_it_init()
      type(suli_ps_opts) :: o
      character(len=64) :: val
      integer :: n, ios
      if (c_associated(suli_ps_it_ctx)) return
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
      suli_ps_it_ctx = suli_ps_create(o)
      if (.not. c_associated(suli_ps_it_ctx)) then
         write (*, *) "suli_ps: no HIP device / context creation failed"
         error stop 1
      end if
   end subroutine suli_ps_it_init

   subroutine suli_ps_it_grid(g)
      use disc
      use cond
      type(suli_ps_grid), intent(out) :: g
      g%nx = nx; g%ny = ny; g%nz = nz
      ! SuLi-IT codes a periodic side as 0, walls otherwise
      g%ccx0 = ccx0; g%ccy0 = ccy0
      g%t_press = 1
      g%obst_t = 0; g%ibm_t = 0
      g%dx = dx; g%dy = dy; g%dz = 0.0d0; g%dt = dt
      g%rho_f1 = 0.0d0
   end subroutine suli_ps_it_grid

   subroutine suli_ps_it_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc < 0) then
         write (*, *) "suli_ps ", where, ": ", trim(suli_ps_error_message(suli_ps_it_ctx))
         error stop 1
      end if
   end subroutine suli_ps_it_check
end module suli_ps_it_state


SUBROUTINE graddin()
   use iso_c_binding
   use dzs
   use velpre
   use parametros
   use suli_ps_iface
   use suli_ps_it_state
   implicit none
   type(suli_ps_grid) :: g
   type(suli_ps_stats) :: st
   integer(c_int) :: rc

   call suli_ps_it_init()
   call suli_ps_it_grid(g)
   rc = suli_ps_it_graddin(suli_ps_it_ctx, g, tetah, gz, u, v, w, dzx, dzy, dzz, eta1, deltah, prd1, st)
   call suli_ps_it_check(rc, "graddin")
END SUBROUTINE graddin

