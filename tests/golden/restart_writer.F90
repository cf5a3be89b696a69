! Writes a SuLi-shaped restart record for the reader's tests: every array
! restart_salva lists (SuLi-IC/11_diversos.f90:593-595), in that order and
! with the shapes SuLi allocates them with (SuLi-IC/0_modules.f90:128-189,
! 273-277, 304-305, 334-337), filled with a known pattern, written by one
! list-directed WRITE as SuLi does.  The same arrays the pressure step reads
! also go, unformatted, to a second file, so that the parse can be checked
! value for value.
!
!   restart_writer NX NY NZ BYXF_FULL OUT_TEXT OUT_RAW
!
! BYXF_FULL = 0 writes byxf with no elements (SuLi never allocates it);
! 1 writes it with the (0:nx1+1, 0:nz+1) shape its one use implies
! (SuLi-IC/3_contorno.f90:170).
program restart_writer
   implicit none
   integer :: nx, ny, nz, nx1, ny1, nz1, full
   character(len=512) :: arg, text_path, raw_path
   real(8), allocatable, dimension(:,:,:) :: u, v, w, ub, vb, wb, prd0, prd1, prd, rho, ls_mu, ls
   real(8), allocatable, dimension(:,:,:) :: hsx, hsy, hsz, hs, kurv, xmu, ymu, zmu, tf_u, tf_v, tf_w
   integer, allocatable, dimension(:,:) :: ku, kv, kw
   real(8), allocatable, dimension(:,:) :: bxx0, bxx1, bxy0, bxz0, bxxf, bxxf1, bxyf, bxzf
   real(8), allocatable, dimension(:,:) :: byx0, byy0, byy1, byz0, byxf, byyf, byyf1, byzf
   real(8), allocatable, dimension(:,:) :: bzx0, bzy0, bzz0, bzz1, bzxf, bzyf, bzzf, bzzf1
   real(8) :: d_max, d_min, ls_m, vol_ini, vol_ins

   call get_command_argument(1, arg); read (arg, *) nx
   call get_command_argument(2, arg); read (arg, *) ny
   call get_command_argument(3, arg); read (arg, *) nz
   call get_command_argument(4, arg); read (arg, *) full
   call get_command_argument(5, text_path)
   call get_command_argument(6, raw_path)
   nx1 = nx + 1; ny1 = ny + 1; nz1 = nz + 1

   allocate(u(0:nx1+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny1+1,0:nz+1), w(0:nx+1,0:ny+1,0:nz1+1))
   allocate(ub(0:nx1+1,0:ny+1,0:nz+1), vb(0:nx+1,0:ny1+1,0:nz+1), wb(0:nx+1,0:ny+1,0:nz1+1))
   allocate(ku(0:nx1+1,0:ny+1), kv(0:nx+1,0:ny1+1), kw(0:nx+1,0:ny+1))
   allocate(prd0(0:nx+1,0:ny+1,0:nz+1), prd1(0:nx+1,0:ny+1,0:nz+1), prd(0:nx+1,0:ny+1,0:nz+1))
   allocate(rho(nx,ny,nz), ls_mu(nx,ny,nz), ls(nx,ny,nz))
   allocate(bxx0(0:ny+1,0:nz+1), bxx1(0:ny+1,0:nz+1), bxy0(0:ny1+1,0:nz+1), bxz0(0:ny+1,0:nz1+1))
   allocate(bxxf(0:ny+1,0:nz+1), bxxf1(0:ny+1,0:nz+1), bxyf(0:ny1+1,0:nz+1), bxzf(0:ny+1,0:nz1+1))
   allocate(byx0(0:nx1+1,0:nz+1), byy0(0:nx+1,0:nz+1), byy1(0:nx+1,0:nz+1), byz0(0:nx+1,0:nz1+1))
   if (full /= 0) then
      allocate(byxf(0:nx1+1,0:nz+1))
   else
      allocate(byxf(0:-1,0:nz+1))
   end if
   allocate(byyf(0:nx+1,0:nz+1), byyf1(0:nx+1,0:nz+1), byzf(0:nx+1,0:nz1+1))
   allocate(bzx0(0:nx1+1,0:ny+1), bzy0(0:nx+1,0:ny1+1), bzz0(0:nx+1,0:ny+1), bzz1(0:nx+1,0:ny+1))
   allocate(bzxf(0:nx1+1,0:ny+1), bzyf(0:nx+1,0:ny1+1), bzzf(0:nx+1,0:ny+1), bzzf1(0:nx+1,0:ny+1))
   allocate(hsx(nx,ny,nz), hsy(nx,ny,nz), hsz(nx,ny,nz), hs(nx,ny,nz), kurv(nx,ny,nz))
   allocate(xmu(nx1,ny,nz), ymu(nx,ny1,nz), zmu(nx,ny,nz1))
   allocate(tf_u(nx1,ny,nz), tf_v(nx,ny1,nz), tf_w(nx,ny,nz1))

   call fill3(u, 1.0d0); call fill3(v, 2.0d0); call fill3(w, 3.0d0)
   call fill3(prd0, 4.0d0); call fill3(prd1, 5.0d0); call fill3(prd, 6.0d0)
   call fill3(rho, 7.0d0); call fill3(ls_mu, 8.0d0); call fill3(ls, 9.0d0)
   call fill3(hs, 10.0d0)
   ! a run of equal values, which list-directed output may write as r*c
   rho(:, :, 1) = 998.0d0
   ! everything else: values the reader must step over
   ub = -1.0d0; vb = -2.0d0; wb = -3.0d0
   ku = 7; kv = -3; kw = 0
   bxx0 = 0.5d0; bxx1 = 0.5d0; bxy0 = 0.25d0; bxz0 = 0.125d0; bxxf = 1.0d-300; bxxf1 = -1.0d300
   bxyf = 3.0d0; bxzf = 3.0d0; byx0 = 1.0d0; byy0 = 1.0d0; byy1 = 1.0d0; byz0 = 1.0d0
   if (full /= 0) byxf = 42.0d0
   byyf = 2.0d0; byyf1 = 2.0d0; byzf = 2.0d0
   bzx0 = 0.0d0; bzy0 = 0.0d0; bzz0 = 0.0d0; bzz1 = 0.0d0; bzxf = 0.0d0; bzyf = 0.0d0
   bzzf = 0.0d0; bzzf1 = 0.0d0
   d_max = 1.5d0; d_min = -1.5d0; ls_m = 0.01d0
   hsx = 0.0d0; hsy = 0.0d0; hsz = 0.0d0; kurv = 1.0d-5
   xmu = 1.0d-3; ymu = 1.0d-3; zmu = 1.0d-3; tf_u = 0.0d0; tf_v = 0.0d0; tf_w = 0.0d0
   vol_ini = 12.0d0; vol_ins = 11.5d0

   open (unit=6661, action='write', file=trim(text_path), status='replace', form='formatted')
   write (6661, *) u, v, w, ku, kv, kw, prd0, prd1, prd, rho, ls_mu, ls, bxx0, bxx1, bxy0, bxz0, bxxf, &
      bxxf1, bxyf, bxzf, byx0, byy0, byy1, byz0, byxf, byyf, byyf1, byzf, bzx0, bzy0, bzz0, bzz1, bzxf, &
      bzyf, bzzf, bzzf1, ub, vb, wb, d_max, d_min, ls_m, hsx, hsy, hsz, hs, kurv, xmu, ymu, zmu, tf_u, &
      tf_v, tf_w, vol_ini, vol_ins
   close (6661)
   open (unit=6662, file=trim(raw_path), access='stream', form='unformatted', status='replace')
   write (6662) u, v, w, prd0, prd1, prd, rho, ls_mu, ls, hs
   close (6662)
contains
   ! a(i) = base + i/7 + sin(i)/3 over the array in storage order
   subroutine fill3(a, base)
      real(8), intent(inout) :: a(:, :, :)
      real(8), intent(in) :: base
      integer :: i1, i2, i3, m
      m = 0
      do i3 = 1, size(a, 3)
         do i2 = 1, size(a, 2)
            do i1 = 1, size(a, 1)
               a(i1, i2, i3) = base + m / 7.0d0 + sin(dble(m)) / 3.0d0
               m = m + 1
            end do
         end do
      end do
   end subroutine fill3
end program restart_writer
