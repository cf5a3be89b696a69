! iso_c_binding interface to libsuli_ps.so (include/suli_ps.h).
!
! Standalone: depends on nothing from SuLi.  suli_ps_graddin.F90 uses it to
! give SuLi drop-in graddin()/posdin() subroutines.
module suli_ps_iface
   use iso_c_binding
   implicit none
   private

   integer(c_int32_t), parameter, public :: SULI_PS_FP64 = 0, SULI_PS_MIXED = 1
   integer(c_int32_t), parameter, public :: SULI_PS_STOP_REFERENCE = 0, SULI_PS_STOP_RELATIVE = 1

   ! struct suli_ps_grid
   type, bind(C), public :: suli_ps_grid
      integer(c_int32_t) :: nx, ny, nz
      integer(c_int32_t) :: ccx0, ccy0
      integer(c_int32_t) :: t_press
      integer(c_int32_t) :: obst_t, ibm_t
      integer(c_int32_t) :: top_bc = 0
      real(c_double) :: dx, dy, dz, dt
      real(c_double) :: rho_f1
   end type suli_ps_grid

   ! struct suli_ps_opts
   type, bind(C), public :: suli_ps_opts
      integer(c_int32_t) :: precision = SULI_PS_FP64
      integer(c_int32_t) :: stop_rule = SULI_PS_STOP_REFERENCE
      real(c_double) :: tol = 1.0d-8
      integer(c_int32_t) :: max_iter = 10000
      integer(c_int32_t) :: chunk = 0
      integer(c_int32_t) :: verbose = 0
      integer(c_int32_t) :: timing = 0
      integer(c_int32_t) :: keep_null = 0
      integer(c_int32_t) :: no_graphs = 0
      integer(c_int32_t) :: autotune = 0
      integer(c_int32_t) :: reuse_staged = 0
      integer(c_int32_t) :: stored_residual = 0
   end type suli_ps_opts

   ! struct suli_ps_stats
   type, bind(C), public :: suli_ps_stats
      integer(c_int32_t) :: status, iters, refreshes, spmv_timed
      real(c_double) :: rr0, bb, rr, thresh, rel_true, solve_ms, spmv_ms
      integer(c_int32_t) :: graph_replays, reserved
   end type suli_ps_stats

   public :: suli_ps_create, suli_ps_destroy, suli_ps_graddin, suli_ps_posdin, suli_ps_last_error
   public :: suli_ps_pressh, suli_ps_posdin_hydro, suli_ps_it_graddin, suli_ps_it_posdin
   public :: suli_ps_it_posdin_layers, suli_ps_exp_press
   public :: suli_ps_error_message

   interface
      function suli_ps_create(opts) bind(C, name="suli_ps_create") result(ctx)
         import :: c_ptr, suli_ps_opts
         type(suli_ps_opts), intent(in) :: opts
         type(c_ptr) :: ctx
      end function suli_ps_create

      subroutine suli_ps_destroy(ctx) bind(C, name="suli_ps_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine suli_ps_destroy

      function suli_ps_last_error(ctx) bind(C, name="suli_ps_last_error") result(msg)
         import :: c_ptr
         type(c_ptr), value :: ctx
         type(c_ptr) :: msg
      end function suli_ps_last_error

      ! Optional dummies of a bind(C) interface are passed as NULL when absent
      ! (obs_ls* off the IBM path, dz/dzz on uniform grids).
      function suli_ps_graddin(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1, st) &
         bind(C, name="suli_ps_graddin") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid, suli_ps_stats
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), intent(in) :: u(*), v(*), w(*), rho(*)
         real(c_double), intent(in), optional :: obs_lsx(*), obs_lsy(*), obs_lsz(*)
         real(c_double), intent(in), optional :: dz(*), dzz(*)
         real(c_double), intent(inout) :: prd1(*)
         type(suli_ps_stats), intent(out) :: st
         integer(c_int) :: rc
      end function suli_ps_graddin

      function suli_ps_posdin(ctx, g, u, v, w, rho, prd1, dz, dzz, prd) &
         bind(C, name="suli_ps_posdin") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         real(c_double), intent(in) :: rho(*), prd1(*)
         real(c_double), intent(in), optional :: dz(*), dzz(*)
         real(c_double), intent(inout), optional :: prd(*)
         integer(c_int) :: rc
      end function suli_ps_posdin

      ! Hydrostatic mode (t_press = 0): pressh and posdin's other branch.
      function suli_ps_pressh(ctx, g, rho, dz, gz, prd1) bind(C, name="suli_ps_pressh") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), intent(in) :: rho(*)
         real(c_double), intent(in), optional :: dz(*)
         real(c_double), value :: gz
         real(c_double), intent(inout) :: prd1(*)
         integer(c_int) :: rc
      end function suli_ps_pressh

      function suli_ps_posdin_hydro(ctx, g, u, v, w, rho, hs, prd1, dz, prd) &
         bind(C, name="suli_ps_posdin_hydro") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         real(c_double), intent(in) :: rho(*), hs(*), prd1(*)
         real(c_double), intent(in), optional :: dz(*)
         real(c_double), intent(inout), optional :: prd(*)
         integer(c_int) :: rc
      end function suli_ps_posdin_hydro

      ! SuLi-IT's non-hydrostatic pressure and velocity correction.
      function suli_ps_it_graddin(ctx, g, tetah, gz, u, v, w, dzx, dzy, dzz, eta1, deltah, prd1, st) &
         bind(C, name="suli_ps_it_graddin") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid, suli_ps_stats
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), value :: tetah, gz
         real(c_double), intent(in) :: u(*), v(*), w(*), dzx(*), dzy(*), dzz(*), eta1(*), deltah(*)
         real(c_double), intent(inout) :: prd1(*)
         type(suli_ps_stats), intent(out) :: st
         integer(c_int) :: rc
      end function suli_ps_it_graddin

      function suli_ps_it_posdin(ctx, g, tetah, gz, u, v, w, dzz, eta1, prd1, prd) &
         bind(C, name="suli_ps_it_posdin") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), value :: tetah, gz
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         real(c_double), intent(in) :: dzz(*), prd1(*)
         real(c_double), intent(inout), optional :: eta1(*), prd(*)
         integer(c_int) :: rc
      end function suli_ps_it_posdin

      function suli_ps_it_posdin_layers(ctx, g, tetah, gz, u, v, w, dz, dzx, dzy, dzz, eta0, eta1, prd1, prd) &
         bind(C, name="suli_ps_it_posdin_layers") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), value :: tetah, gz
         real(c_double), intent(inout) :: u(*), v(*), w(*), dz(*), dzx(*), dzy(*), dzz(*), eta1(*)
         real(c_double), intent(in) :: eta0(*), prd1(*)
         real(c_double), intent(inout), optional :: prd(*)
         integer(c_int) :: rc
      end function suli_ps_it_posdin_layers

      function suli_ps_exp_press(ctx, g, rho, prd1, prd0, dz, dzz, px, py, pz) &
         bind(C, name="suli_ps_exp_press") result(rc)
         import :: c_ptr, c_double, c_int, suli_ps_grid
         type(c_ptr), value :: ctx
         type(suli_ps_grid), intent(in) :: g
         real(c_double), intent(in) :: rho(*), prd1(*)
         real(c_double), intent(inout) :: prd0(*)
         real(c_double), intent(in), optional :: dz(*), dzz(*)
         real(c_double), intent(out) :: px(*), py(*), pz(*)
         integer(c_int) :: rc
      end function suli_ps_exp_press
   end interface

contains

   ! Copy of the library's last error message as a Fortran string.
   function suli_ps_error_message(ctx) result(text)
      type(c_ptr), intent(in) :: ctx
      character(len=512) :: text
      character(kind=c_char), pointer :: chars(:)
      type(c_ptr) :: p
      integer :: i
      text = ' '
      p = suli_ps_last_error(ctx)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, chars, [512])
      do i = 1, 512
         if (chars(i) == c_null_char) exit
         text(i:i) = chars(i)
      end do
   end function suli_ps_error_message

end module suli_ps_iface
