! Drop-in posdin() for SuLi-IT (SuLi-IT/10_pos.f90:73-132) on the MI355X
! library; the context and helpers are suli_ps_graddin_it.F90's.  All of
! it runs on the GPU in the reference's order: the free surface and its
! layer (eta1, dz(:,:,nz)), the top planes' face thicknesses (dz_it,
! SuLi-IT/11_diversos.f90:80-141), then the velocity correction with the
! new dzz.
SUBROUTINE posdin()
   use iso_c_binding
   use disc
   use dzs
   use velpre
   use parametros
   use suli_ps_iface
   use suli_ps_it_state
   implicit none
   type(suli_ps_grid) :: g
   integer(c_int) :: rc

   call suli_ps_it_init()
   call suli_ps_it_grid(g)
   rc = suli_ps_it_posdin_layers(suli_ps_it_ctx, g, tetah, gz, u, v, w, dz, dzx, dzy, dzz, eta0, eta1, prd1, prd)
   call suli_ps_it_check(rc, "posdin")
END SUBROUTINE posdin
