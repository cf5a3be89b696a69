! End-to-end check of the drop-in: a SuLi-style program that fills SuLi's
! own module arrays, calls graddin()/posdin() (suli_ps_graddin.F90 -> GPU),
! then calls the reference's graddin/posdin (linked in under the names
! ref_graddin/ref_posdin by oracle/Makefile) on the same inputs and compares.
! MODE 1 is the hydrostatic step instead (t_press = 0,
! SuLi-IC/1_principal.f90:97-100): pressh() then posdin(), against
! ref_pressh/ref_posdin.  MODE 2 is the IPA step (t_press = 2, :82 and
! :111): exp_press() then graddin(), against ref_exp_press/ref_graddin.
!
!   dropin_ic NX NY NZ PERIODIC_X MIXED [MODE]
!
! Prints the maximum relative differences and exits with status 0 when the
! pressure agrees to 1e-8 (interior and face ghosts) and the corrected
! velocities to 1e-8 relative (hydrostatic step: 1e-14, the library's
! column sweeps repeat the reference's arithmetic); IPA: prd0, px, py and
! pz from exp_press to 1e-15 relative (the same IEEE operations in the same
! order), the pressure to 1e-8.
program dropin_check
   use disc
   use cond
   use velpre
   use obst
   use ls_param
   use param, only: gz
   implicit none
   integer :: i, j, k, argn(6), nargs
   character(len=32) :: arg
   real(8) :: xc, zc, hgt, amp, dp, scale, du, tol, dq
   real(8), allocatable :: u0(:, :, :), v0(:, :, :), w0(:, :, :)
   real(8), allocatable :: pg(:, :, :), ug(:, :, :), vg(:, :, :), wg(:, :, :)
   real(8), allocatable :: p1s(:, :, :), p0s(:, :, :), p0g(:, :, :), pxg(:, :, :), pyg(:, :, :), pzg(:, :, :)
   logical :: ok, hydro, ipa
   external :: ref_graddin, ref_posdin, ref_pressh, ref_exp_press

   argn = [24, 16, 12, 0, 0, 0]
   nargs = command_argument_count()
   do i = 1, min(nargs, 6)
      call get_command_argument(i, arg)
      read (arg, *) argn(i)
   end do
   nx = argn(1); ny = argn(2); nz = argn(3)
   nx1 = nx + 1; ny1 = ny + 1; nz1 = nz + 1
   dx = 5.0d-3; dy = 5.0d-3; dt = 1.0d-3
#ifdef SULI_ICDZ
   allocate(dz(0:nz+1), dzz(0:nz+1))
   do k = 0, nz + 1
      dz(k) = 5.0d-3 * (1.0d0 + 0.02d0 * k)
   end do
   do k = 1, nz + 1
      dzz(k) = (dz(k) + dz(k-1)) * 0.5d0
   end do
   dzz(0) = dzz(1)
#else
   dz = 5.0d-3
#endif
   ccx0 = 1; ccxf = 1; ccy0 = 1; ccyf = 1
   if (argn(4) /= 0) then
      ccx0 = 0; ccxf = 0
   end if
   hydro = argn(6) == 1
   ipa = argn(6) == 2
   t_press = 1; obst_t = 0; ibm_t = 0; omp_t = 0
   if (hydro) t_press = 0
   if (ipa) t_press = 2
   gz = 9.80665d0
   rho_f1 = 1.2d0; rho_f2 = 1000.0d0

   allocate(u(0:nx1+1, 0:ny+1, 0:nz+1), v(0:nx+1, 0:ny1+1, 0:nz+1), w(0:nx+1, 0:ny+1, 0:nz1+1))
   allocate(prd1(0:nx+1, 0:ny+1, 0:nz+1), prd(0:nx+1, 0:ny+1, 0:nz+1), rho(nx, ny, nz))
   allocate(obs_lsx(nx1, ny, nz), obs_lsy(nx, ny1, nz), obs_lsz(nx, ny, nz1))

   ! two-phase density with a smooth interface at mid depth
   hgt = 0.5d0 * nz * 5.0d-3
   do k = 1, nz
      do j = 1, ny
         do i = 1, nx
            xc = (i - 0.5d0) * dx
            zc = (k - 0.5d0) * 5.0d-3
            rho(i, j, k) = rho_f1 + (rho_f2 - rho_f1) * 0.5d0 * &
                           (1.0d0 - tanh((zc - hgt - 0.01d0 * sin(6.2831853d0 * xc / (nx * dx))) / 0.0075d0))
         end do
      end do
   end do
   ! deterministic divergent velocity on interior faces, zero normal walls;
   ! large amplitude so the reference's absolute stop rule is a tight one
   amp = 2.0d3
   u = 0.0d0; v = 0.0d0; w = 0.0d0
   do k = 1, nz
      do j = 1, ny
         do i = 2, nx
            u(i, j, k) = amp * sin(0.7d0 * i + 1.3d0 * j) * cos(0.9d0 * k)
         end do
         if (ccx0 == 0) then
            u(1, j, k) = amp * sin(0.7d0 + 1.3d0 * j) * cos(0.9d0 * k)
            u(nx + 1, j, k) = u(1, j, k)
         end if
         do i = 1, nx
            if (j >= 2) v(i, j, k) = amp * cos(0.5d0 * i - 0.8d0 * j) * sin(1.1d0 * k)
#ifndef SULI_ICDZ
            ! (on a SuLi-ICdz grid a w field must satisfy a per-column
            ! compatibility condition; u and v alone keep the system solvable)
            if (k >= 2) w(i, j, k) = amp * sin(0.3d0 * i * j + 0.4d0 * k)
#endif
         end do
      end do
   end do
   prd1 = 0.0d0
   u0 = u; v0 = v; w0 = w
   if (ipa) then
      ! the pressures a previous step left: prd0 accumulated, prd1 its last
      ! increment (also graddin's initial guess)
      allocate(prd0(0:nx+1, 0:ny+1, 0:nz+1), px(nx1, ny, nz), py(nx, ny1, nz), pz(nx, ny, nz1))
      do k = 0, nz + 1
         do j = 0, ny + 1
            do i = 0, nx + 1
               prd1(i, j, k) = 1.0d2 * sin(0.3d0 * i + 0.2d0 * j) * cos(0.25d0 * k)
               prd0(i, j, k) = 1.0d3 * cos(0.1d0 * i - 0.15d0 * j + 0.05d0 * k)
            end do
         end do
      end do
      ! prd1's ghosts as graddin leaves them (x, then y, then z: mirrored,
      ! or wrapped in a periodic x; SuLi-IC/9_grad.f90:594-643): a warm
      ! start whose ghosts do not follow the interior gives the reference an
      ! inconsistent singular system, whose iterate drifts along the null
      ! space
      if (ccx0 == 0) then
         prd1(0, :, :) = prd1(nx, :, :); prd1(nx + 1, :, :) = prd1(1, :, :)
      else
         prd1(0, :, :) = prd1(1, :, :); prd1(nx + 1, :, :) = prd1(nx, :, :)
      end if
      prd1(:, 0, :) = prd1(:, 1, :); prd1(:, ny + 1, :) = prd1(:, ny, :)
      prd1(:, :, 0) = prd1(:, :, 1); prd1(:, :, nz + 1) = prd1(:, :, nz)
      p1s = prd1; p0s = prd0
      call exp_press()
      p0g = prd0; pxg = px; pyg = py; pzg = pz
      call graddin()
   else if (hydro) then
      ! the Heaviside field of the same interface: exactly 0 well above it
      allocate(hs(nx, ny, nz))
      hs = (rho - rho_f1) / (rho_f2 - rho_f1)
      where (hs < 1.0d-3) hs = 0.0d0
      call pressh()
      call posdin()
   else
      call graddin()
      call posdin()
   end if
   pg = prd1; ug = u; vg = v; wg = w

   u = u0; v = v0; w = w0; prd1 = 0.0d0
   if (ipa) then
      prd1 = p1s; prd0 = p0s
      call ref_exp_press()
      call ref_graddin()
   else if (hydro) then
      call ref_pressh()
      call ref_posdin()
   else
      call ref_graddin()
      call ref_posdin()
   end if

   scale = maxval(abs(prd1(1:nx, 1:ny, 1:nz)))
   dp = 0.0d0
   do k = 0, nz + 1
      do j = 0, ny + 1
         do i = 0, nx + 1
            ! interior and face ghosts only (edges/corners race in the reference)
            if (count([i == 0 .or. i == nx + 1, j == 0 .or. j == ny + 1, k == 0 .or. k == nz + 1]) <= 1) &
               dp = max(dp, abs(pg(i, j, k) - prd1(i, j, k)))
         end do
      end do
   end do
   dp = dp / scale
   if (ipa) then
      ! no posdin in the IPA step: the velocities are untouched on both sides
      du = max(maxval(abs(ug - u)), maxval(abs(vg - v)), maxval(abs(wg - w)))
      dq = max(maxval(abs(p0g - prd0)) / maxval(abs(prd0)), maxval(abs(pxg - px)) / maxval(abs(px)), &
               maxval(abs(pyg - py)) / maxval(abs(py)), maxval(abs(pzg - pz)) / maxval(abs(pz)))
   else
      du = max(maxval(abs(ug - u)) / maxval(abs(u)), maxval(abs(vg - v)) / maxval(abs(v)), &
               maxval(abs(wg - w)) / maxval(abs(w)))
      dq = 0.0d0
   end if
   tol = 1.0d-8
   if (hydro) tol = 1.0d-14
   ok = dp <= tol .and. du <= tol .and. dq <= 1.0d-15
   if (ipa) write (*, '(a,es10.3)') "exp_press max rel diff =", dq
   write (*, '(a,3i5,a,es10.3,a,es10.3,a,l2)') "dropin", nx, ny, nz, "  max rel dp =", dp, &
      "  max rel du =", du, "  ok =", ok
   if (.not. ok) error stop 1
end program dropin_check
