! End-to-end check of the SuLi-IT drop-in: a SuLi-IT-style program fills
! SuLi-IT's own module arrays (their size is SuLi-IT's compile-time grid),
! calls graddin() (suli_ps_graddin_it.F90 -> GPU), then the reference's
! graddin (linked in as ref_graddin by oracle/Makefile) on the same inputs,
! and compares the pressure.  The velocities are large enough that
! SuLi-IT's absolute stop rule (|r.r| <= 1e-11) is a tight one.
!
!   dropin_it [MIXED]        (the precision comes from SULI_PS_PRECISION)
!
! Exits with status 0 when prd1 agrees to 1e-8 relative.
program dropin_check_it
   use disc
   use velpre
   use dzs
   use parametros
   implicit none
   integer :: i, j, k
   real(8) :: amp, dp
   real(8), allocatable :: pg(:, :, :)
   logical :: ok
   external :: ref_graddin

   amp = 1.0d3
   do k = 0, nz + 1
      do j = 0, ny + 1
         do i = 0, nx1 + 1
            dzx(i, j, k) = dz1 * (1.0d0 + 0.2d0 * sin(0.3d0 * i + 0.7d0 * j + 1.1d0 * k))
         end do
         do i = 0, nx + 1
            dzy(i, j, k) = dz1 * (1.0d0 + 0.2d0 * cos(0.5d0 * i - 0.4d0 * j + 0.9d0 * k))
         end do
      end do
   end do
   do k = 0, nz1 + 1
      do j = 0, ny + 1
         do i = 0, nx + 1
            dzz(i, j, k) = dz1 * (1.0d0 + 0.15d0 * sin(0.2d0 * i * j + 0.6d0 * k))
         end do
      end do
   end do
   u = 0.0d0; v = 0.0d0; w = 0.0d0
   do k = 1, nz
      do j = 1, ny
         do i = 1, nx1
            u(i, j, k) = amp * 0.1d0 * sin(0.7d0 * i + 1.3d0 * j) * cos(0.9d0 * k)
         end do
      end do
      do j = 1, ny1
         do i = 1, nx
            v(i, j, k) = amp * 0.1d0 * cos(0.5d0 * i - 0.8d0 * j) * sin(1.1d0 * k)
         end do
      end do
   end do
   do k = 1, nz1
      do j = 1, ny
         do i = 1, nx
            w(i, j, k) = amp * 0.02d0 * sin(0.3d0 * i * j + 0.4d0 * k)
         end do
      end do
   end do
   do j = 0, ny + 1
      do i = 0, nx + 1
         eta1(i, j) = amp * 1.0d-5 * sin(0.2d0 * i + 0.3d0 * j)
      end do
   end do
   do j = 1, ny
      do i = 1, nx
         deltah(i, j) = amp * 1.0d-5 * cos(0.25d0 * i - 0.35d0 * j)
      end do
   end do

   prd1 = 0.0d0
   call graddin()
   allocate(pg(0:nx + 1, 0:ny + 1, 0:nz + 1))
   pg = prd1
   prd1 = 0.0d0
   call ref_graddin()
   dp = maxval(abs(pg - prd1)) / maxval(abs(prd1(1:nx, 1:ny, 1:nz)))
   ok = dp <= 1.0d-8
   write (*, '(a,3i5,a,es10.3,a,l2)') "dropin_it", nx, ny, nz, "  max rel dp =", dp, "  ok =", ok
   if (.not. ok) error stop 1
end program dropin_check_it
