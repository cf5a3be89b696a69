! Flat C interface to the reference's SuLi-IT graddin (TEST INFRASTRUCTURE:
! the checker for the GPU path, never part of the product).  SuLi-IT fixes
! its grid at compile time (SuLi-IT/0_modules.f90:8-15: 50 x 50 x 25) and
! its walls (cond, :57-66: ccx0 = ccy0 = 2), so the bridge exposes those
! values and copies arrays of exactly that shape in and out of the modules.
module suli_ref_bridge_it
   use iso_c_binding
   implicit none
contains
   ! dims(3) = nx, ny, nz; par(5) = dx, dy, dt, tetah, gz
   subroutine ref_it_info(dims, par) bind(C, name="suli_ref_it_info")
      use disc
      use parametros
      integer(c_int), intent(out) :: dims(3)
      real(c_double), intent(out) :: par(5)
      dims = [nx, ny, nz]
      par = [dx, dy, dt, tetah, gz]
   end subroutine ref_it_info

   ! One graddin() on the given state; prd1 in (initial guess) and out.
   subroutine ref_it_graddin(uin, vin, win, dzxin, dzyin, dzzin, eta1in, deltahin, prd1io) &
      bind(C, name="suli_ref_it_graddin")
      use disc
      use velpre
      use dzs
      real(c_double), intent(in) :: uin(0:nx1+1, 0:ny+1, 0:nz+1), dzxin(0:nx1+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(in) :: vin(0:nx+1, 0:ny1+1, 0:nz+1), dzyin(0:nx+1, 0:ny1+1, 0:nz+1)
      real(c_double), intent(in) :: win(0:nx+1, 0:ny+1, 0:nz1+1), dzzin(0:nx+1, 0:ny+1, 0:nz1+1)
      real(c_double), intent(in) :: eta1in(0:nx+1, 0:ny+1), deltahin(nx, ny)
      real(c_double), intent(inout) :: prd1io(0:nx+1, 0:ny+1, 0:nz+1)
      u = uin; v = vin; w = win
      dzx = dzxin; dzy = dzyin; dzz = dzzin
      eta1 = eta1in; deltah = deltahin
      prd1 = prd1io
      call graddin()
      prd1io = prd1
   end subroutine ref_it_graddin

   ! One posdin() (SuLi-IT/10_pos.f90:73-132 with dz_it, compiled on its own
   ! by oracle/Makefile): u, v, w, dz, dzx, dzy, dzz and eta1 in and out,
   ! eta0 and prd1 in, prd out.
   subroutine ref_it_posdin(uio, vio, wio, dzio, dzxio, dzyio, dzzio, eta0in, eta1io, prd1in, prdout) &
      bind(C, name="suli_ref_it_posdin")
      use disc
      use velpre
      use dzs
      real(c_double), intent(inout) :: uio(0:nx1+1, 0:ny+1, 0:nz+1), dzxio(0:nx1+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(inout) :: vio(0:nx+1, 0:ny1+1, 0:nz+1), dzyio(0:nx+1, 0:ny1+1, 0:nz+1)
      real(c_double), intent(inout) :: wio(0:nx+1, 0:ny+1, 0:nz1+1), dzzio(0:nx+1, 0:ny+1, 0:nz1+1)
      real(c_double), intent(inout) :: dzio(0:nx+1, 0:ny+1, 0:nz+1), eta1io(0:nx+1, 0:ny+1)
      real(c_double), intent(in) :: eta0in(0:nx+1, 0:ny+1), prd1in(0:nx+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(out) :: prdout(0:nx+1, 0:ny+1, 0:nz+1)
      u = uio; v = vio; w = wio
      dz = dzio; dzx = dzxio; dzy = dzyio; dzz = dzzio
      eta0 = eta0in; eta1 = eta1io
      prd1 = prd1in
      call posdin()
      uio = u; vio = v; wio = w
      dzio = dz; dzxio = dzx; dzyio = dzy; dzzio = dzz
      eta1io = eta1
      prdout = prd
   end subroutine ref_it_posdin
end module suli_ref_bridge_it
