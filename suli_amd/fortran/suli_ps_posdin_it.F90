! Drop-in posdin() for SuLi-IT (SuLi-IT/10_pos.f90:73-132) on the MI355X
! library; the context and helpers are suli_ps_graddin_it.F90's.  The free
! surface and its layer (eta1, dz(:,:,nz), dz_it) stay on the host, as
! SuLi-IT keeps them; the velocity correction runs on the GPU.
SUBROUTINE posdin()
   use iso_c_binding
   use dzs
   use velpre
   use parametros
   use suli_ps_iface
   use suli_ps_it_state
   implicit none
   type(suli_ps_grid) :: g
   integer(c_int) :: rc
   integer :: i, j

   call suli_ps_it_init()
   call suli_ps_it_grid(g)
   ! the free surface and its layer first (SuLi-IT/10_pos.f90:84-93), then
   ! the correction with the updated dzz, as the reference orders them
   do j = 0, ny + 1
      do i = 0, nx + 1
         eta1(i, j) = eta1(i, j) + prd1(i, j, nz) / gz
         dz(i, j, nz) = dz(i, j, nz) + (eta1(i, j) - eta0(i, j))
      end do
   end do
   call dz_it()
   rc = suli_ps_it_posdin(suli_ps_it_ctx, g, tetah, gz, u, v, w, dzz, prd1=prd1, prd=prd)
   call suli_ps_it_check(rc, "posdin")
END SUBROUTINE posdin
