! End-to-end check of the SuLi-IT drop-in: a SuLi-IT-style program fills
! SuLi-IT's own module arrays (their size is SuLi-IT's compile-time grid),
! calls graddin() (suli_ps_graddin_it.F90 -> GPU), then the reference's
! graddin (linked in as ref_graddin by oracle/Makefile) on the same inputs,
! and compares the pressure.  The velocities are large enough that
! SuLi-IT's absolute stop rule (|r.r| <= 1e-11) is a tight one.  Then
! posdin() (suli_ps_posdin_it.F90 -> GPU: surface, top layer, dz_it,
! correction) against ref_posdin on the same state and pressure.
!
!   dropin_it                (the precision comes from SULI_PS_PRECISION)
!
! Exits with status 0 when prd1 agrees to 1e-8 relative and everything
! posdin writes to 1e-15 relative (the same IEEE operations in order).
program dropin_check_it
   use disc
   use velpre
   use dzs
   use parametros
   implicit none
   integer :: i, j, k
   real(8) :: amp, dp, dq
   real(8), allocatable :: pg(:, :, :)
   real(8), allocatable :: s_u(:, :, :), s_v(:, :, :), s_w(:, :, :), s_dz(:, :, :), s_dzx(:, :, :)
   real(8), allocatable :: s_dzy(:, :, :), s_dzz(:, :, :), s_eta1(:, :)
   real(8), allocatable :: g_u(:, :, :), g_v(:, :, :), g_w(:, :, :), g_dz(:, :, :), g_dzx(:, :, :)
   real(8), allocatable :: g_dzy(:, :, :), g_dzz(:, :, :), g_eta1(:, :)
   logical :: ok
   external :: ref_graddin, ref_posdin

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

   ! posdin on the GPU's pressure, both sides from the same state: the
   ! surface a little above eta0, cell thicknesses about dz1
   do k = 0, nz + 1
      do j = 0, ny + 1
         do i = 0, nx + 1
            dz(i, j, k) = dz1 * (1.0d0 + 0.1d0 * sin(0.4d0 * i - 0.3d0 * j + 0.7d0 * k))
         end do
      end do
   end do
   eta0 = eta1 - amp * 2.0d-6
   prd1 = pg
   s_u = u; s_v = v; s_w = w; s_dz = dz; s_dzx = dzx; s_dzy = dzy; s_dzz = dzz; s_eta1 = eta1
   call posdin()
   g_u = u; g_v = v; g_w = w; g_dz = dz; g_dzx = dzx; g_dzy = dzy; g_dzz = dzz; g_eta1 = eta1
   u = s_u; v = s_v; w = s_w; dz = s_dz; dzx = s_dzx; dzy = s_dzy; dzz = s_dzz; eta1 = s_eta1
   prd1 = pg
   call ref_posdin()
   dq = max(maxval(abs(g_u - u)) / maxval(abs(u)), maxval(abs(g_v - v)) / maxval(abs(v)), &
            maxval(abs(g_w - w)) / maxval(abs(w)), maxval(abs(g_dz - dz)) / maxval(abs(dz)), &
            maxval(abs(g_dzx - dzx)) / maxval(abs(dzx)), maxval(abs(g_dzy - dzy)) / maxval(abs(dzy)), &
            maxval(abs(g_dzz - dzz)) / maxval(abs(dzz)), maxval(abs(g_eta1 - eta1)) / maxval(abs(eta1)))
   ok = dp <= 1.0d-8 .and. dq <= 1.0d-15
   write (*, '(a,3i5,a,es10.3,a,es10.3,a,l2)') "dropin_it", nx, ny, nz, "  max rel dp =", dp, &
      "  posdin max rel diff =", dq, "  ok =", ok
   if (.not. ok) error stop 1
end program dropin_check_it
