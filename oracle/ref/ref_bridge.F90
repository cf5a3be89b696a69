! Test-infrastructure bridge to the reference SuLi pressure step.
!
! This file is NOT part of the product. It is compiled (by oracle/Makefile,
! in the development container only) together with the reference's own
! hot-path sources
!     SuLi-IC/0_modules.f90, 8_derivadas.f90, 9_grad.f90, 10_pos.f90
!   or SuLi-ICdz/ (same files, built with -DSULI_ICDZ)
! into oracle/_ref/libsuli_ref_ic.so / libsuli_ref_icdz.so.  It exposes a
! flat C ABI so that the tests can run the reference graddin()/posdin()
! (SuLi-IC/9_grad.f90:10-663, SuLi-IC/10_pos.f90:1-99) on identical inputs.
!
! The reference's subroutines take no arguments: everything flows through
! module variables (SuLi-IC/0_modules.f90:2-58 disc, :68-80 cond,
! :108-118 obst, :144-162 velpre, :282-295 ls_param).  The bridge sets those
! module variables, copies the caller's arrays in, calls the reference, and
! copies the results out.  All arrays use the reference's Fortran layouts
! (column-major, x fastest, halos as declared at 0_modules.f90:170-172,188-189).
module suli_ref_bridge
   use iso_c_binding
   implicit none
contains

   subroutine ref_setup(nx_, ny_, nz_, dx_, dy_, dz_, dt_, ccx0_, ccy0_, t_press_, &
                        obst_t_, ibm_t_, omp_t_, rho_f1_, dzarr, dzzarr) &
      bind(C, name="suli_ref_setup")
      use disc
      use cond
      use velpre
      use obst
      use ls_param
      integer(c_int), value :: nx_, ny_, nz_, ccx0_, ccy0_, t_press_, obst_t_, ibm_t_, omp_t_
      real(c_double), value :: dx_, dy_, dz_, dt_, rho_f1_
      ! dz(0:nz+1) and dzz(0:nz+1); read only in the ICdz build
      real(c_double), intent(in) :: dzarr(0:nz_+1), dzzarr(0:nz_+1)

      nx = nx_; ny = ny_; nz = nz_
      nx1 = nx + 1; ny1 = ny + 1; nz1 = nz + 1
      dx = dx_; dy = dy_
      dt = dt_
      ccx0 = ccx0_; ccxf = ccx0_
      ccy0 = ccy0_; ccyf = ccy0_
      t_press = t_press_
      obst_t = obst_t_; ibm_t = ibm_t_
      omp_t = omp_t_
      rho_f1 = rho_f1_

#ifdef SULI_ICDZ
      if (allocated(dz)) deallocate(dz)
      if (allocated(dzz)) deallocate(dzz)
      allocate(dz(0:nz+1), dzz(0:nz+1))
      dz = dzarr
      dzz = dzzarr
#else
      dz = dz_
      if (.false.) dz = dzarr(0) + dzzarr(0)
#endif

      if (allocated(u)) deallocate(u, v, w, prd1, prd, rho)
      allocate(u(0:nx1+1, 0:ny+1, 0:nz+1))
      allocate(v(0:nx+1, 0:ny1+1, 0:nz+1))
      allocate(w(0:nx+1, 0:ny+1, 0:nz1+1))
      allocate(prd1(0:nx+1, 0:ny+1, 0:nz+1), prd(0:nx+1, 0:ny+1, 0:nz+1))
      allocate(rho(nx, ny, nz))
      if (allocated(obs_lsx)) deallocate(obs_lsx, obs_lsy, obs_lsz)
      allocate(obs_lsx(nx1, ny, nz), obs_lsy(nx, ny1, nz), obs_lsz(nx, ny, nz1))
      obs_lsx = -1.0d0
      obs_lsy = -1.0d0
      obs_lsz = -1.0d0
   end subroutine ref_setup

   ! Runs the reference graddin() once.  u/v/w/rho/prd1 are read in the
   ! reference layouts; prd1 (the full halo box) is written back.  When
   ! have_obs /= 0 the three obs_ls* face arrays are copied in as well.
   subroutine ref_graddin(uin, vin, win, rhoin, have_obs, obsx, obsy, obsz, prd1io) &
      bind(C, name="suli_ref_graddin")
      use disc
      use velpre
      use obst
      real(c_double), intent(in) :: uin(0:nx+2, 0:ny+1, 0:nz+1)
      real(c_double), intent(in) :: vin(0:nx+1, 0:ny+2, 0:nz+1)
      real(c_double), intent(in) :: win(0:nx+1, 0:ny+1, 0:nz+2)
      real(c_double), intent(in) :: rhoin(nx, ny, nz)
      integer(c_int), value :: have_obs
      real(c_double), intent(in) :: obsx(nx+1, ny, nz), obsy(nx, ny+1, nz), obsz(nx, ny, nz+1)
      real(c_double), intent(inout) :: prd1io(0:nx+1, 0:ny+1, 0:nz+1)

      u = uin
      v = vin
      w = win
      rho = rhoin
      if (have_obs /= 0) then
         obs_lsx = obsx
         obs_lsy = obsy
         obs_lsz = obsz
      end if
      prd1 = prd1io
      call graddin()
      prd1io = prd1
   end subroutine ref_graddin

   ! Runs the reference posdin() once: u/v/w are updated in place and prd
   ! is written (posdin copies prd = prd1).
   subroutine ref_posdin(uio, vio, wio, rhoin, prd1in, prdout) bind(C, name="suli_ref_posdin")
      use disc
      use velpre
      real(c_double), intent(inout) :: uio(0:nx+2, 0:ny+1, 0:nz+1)
      real(c_double), intent(inout) :: vio(0:nx+1, 0:ny+2, 0:nz+1)
      real(c_double), intent(inout) :: wio(0:nx+1, 0:ny+1, 0:nz+2)
      real(c_double), intent(in) :: rhoin(nx, ny, nz)
      real(c_double), intent(in) :: prd1in(0:nx+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(out) :: prdout(0:nx+1, 0:ny+1, 0:nz+1)

      u = uio
      v = vio
      w = wio
      rho = rhoin
      prd1 = prd1in
      call posdin()
      uio = u
      vio = v
      wio = w
      prdout = prd
   end subroutine ref_posdin

   ! Runs the reference pressh() (hydrostatic pressure, t_press = 0) once:
   ! prd1's interior is written from rho and gz; its halo is left as given.
   subroutine ref_pressh(rhoin, gz_, prd1io) bind(C, name="suli_ref_pressh")
      use disc
      use velpre
      use param
      real(c_double), intent(in) :: rhoin(nx, ny, nz)
      real(c_double), value :: gz_
      real(c_double), intent(inout) :: prd1io(0:nx+1, 0:ny+1, 0:nz+1)

      rho = rhoin
      gz = gz_
      prd1 = prd1io
      call pressh()
      prd1io = prd1
   end subroutine ref_pressh

   ! Runs the reference posdin() in its hydrostatic branch (t_press /= 1,
   ! set up by suli_ref_setup): u, v corrected from prd1 and rho, w from
   ! continuity, velocities zeroed where the face value of hs is below 1e-4,
   ! prd = prd1.
   subroutine ref_posdin_hydro(uio, vio, wio, rhoin, hsin, prd1in, prdout) bind(C, name="suli_ref_posdin_hydro")
      use disc
      use velpre
      use ls_param
      real(c_double), intent(inout) :: uio(0:nx+2, 0:ny+1, 0:nz+1)
      real(c_double), intent(inout) :: vio(0:nx+1, 0:ny+2, 0:nz+1)
      real(c_double), intent(inout) :: wio(0:nx+1, 0:ny+1, 0:nz+2)
      real(c_double), intent(in) :: rhoin(nx, ny, nz)
      real(c_double), intent(in) :: hsin(nx, ny, nz)
      real(c_double), intent(in) :: prd1in(0:nx+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(out) :: prdout(0:nx+1, 0:ny+1, 0:nz+1)

      if (allocated(hs)) deallocate(hs)
      allocate(hs(nx, ny, nz))
      hs = hsin
      u = uio
      v = vio
      w = wio
      rho = rhoin
      prd1 = prd1in
      call posdin()
      uio = u
      vio = v
      wio = w
      prdout = prd
   end subroutine ref_posdin_hydro

   ! Runs the reference exp_press() (SuLi-IC/11_diversos.f90:455-495, or
   ! SuLi-ICdz's, compiled on its own by oracle/Makefile) on the grid that
   ! ref_setup fixed: prd0 is updated in place, px, py, pz are written.
   subroutine ref_exp_press(rhoin, prd1in, prd0io, pxout, pyout, pzout) &
      bind(C, name="suli_ref_exp_press")
      use disc
      use velpre
      real(c_double), intent(in) :: rhoin(nx, ny, nz), prd1in(0:nx+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(inout) :: prd0io(0:nx+1, 0:ny+1, 0:nz+1)
      real(c_double), intent(out) :: pxout(nx+1, ny, nz), pyout(nx, ny+1, nz), pzout(nx, ny, nz+1)

      if (allocated(prd0)) deallocate(prd0)
      if (allocated(px)) deallocate(px)
      if (allocated(py)) deallocate(py)
      if (allocated(pz)) deallocate(pz)
      allocate(prd0(0:nx+1, 0:ny+1, 0:nz+1), px(nx1, ny, nz), py(nx, ny1, nz), pz(nx, ny, nz1))
      rho = rhoin
      prd1 = prd1in
      prd0 = prd0io
      call exp_press()
      prd0io = prd0
      pxout = px
      pyout = py
      pzout = pz
   end subroutine ref_exp_press

#ifndef SULI_ICDZ
   ! Runs the reference heaviside() (SuLi-IC/14_level_set.f90:508-594) with
   ! t_tens = 0: hs, rho and ls_mu from the level set ls, alpha1 and dx1.
   subroutine ref_heaviside(lsin, alpha1_, dx1_, rho1, rho2, mu1, mu2, hsout, rhoout, muout) &
      bind(C, name="suli_ref_heaviside")
      use disc
      use velpre
      use ls_param
      real(c_double), intent(in) :: lsin(nx, ny, nz)
      real(c_double), value :: alpha1_, dx1_, rho1, rho2, mu1, mu2
      real(c_double), intent(out) :: hsout(nx, ny, nz), rhoout(nx, ny, nz), muout(nx, ny, nz)

      if (allocated(ls)) deallocate(ls)
      if (allocated(hs)) deallocate(hs)
      if (allocated(ls_mu)) deallocate(ls_mu)
      if (allocated(rho)) deallocate(rho)
      allocate(ls(nx, ny, nz), hs(nx, ny, nz), ls_mu(nx, ny, nz), rho(nx, ny, nz))
      ls = lsin
      t_tens = 0
      alpha1 = alpha1_
      dx1 = dx1_
      rho_f1 = rho1; rho_f2 = rho2
      mu_f1 = mu1; mu_f2 = mu2
      call heaviside()
      hsout = hs
      rhoout = rho
      muout = ls_mu
   end subroutine ref_heaviside

   ! heaviside() with t_tens = 1 (:519-550, 563-567): also the surface-tension
   ! gradients hsx, hsy, hsz, from der_weno on the grid ref_setup fixed
   ! (dx, dy, dz, ccx0, ccy0).
   subroutine ref_heaviside_tens(lsin, alpha1_, dx1_, rho1, rho2, mu1, mu2, hsout, rhoout, muout, &
                                 hsxout, hsyout, hszout) bind(C, name="suli_ref_heaviside_tens")
      use disc
      use velpre
      use ls_param
      real(c_double), intent(in) :: lsin(nx, ny, nz)
      real(c_double), value :: alpha1_, dx1_, rho1, rho2, mu1, mu2
      real(c_double), intent(out) :: hsout(nx, ny, nz), rhoout(nx, ny, nz), muout(nx, ny, nz)
      real(c_double), intent(out) :: hsxout(nx, ny, nz), hsyout(nx, ny, nz), hszout(nx, ny, nz)

      if (allocated(ls)) deallocate(ls)
      if (allocated(hs)) deallocate(hs)
      if (allocated(ls_mu)) deallocate(ls_mu)
      if (allocated(rho)) deallocate(rho)
      if (allocated(hsx)) deallocate(hsx)
      if (allocated(hsy)) deallocate(hsy)
      if (allocated(hsz)) deallocate(hsz)
      allocate(ls(nx, ny, nz), hs(nx, ny, nz), ls_mu(nx, ny, nz), rho(nx, ny, nz))
      allocate(hsx(nx, ny, nz), hsy(nx, ny, nz), hsz(nx, ny, nz))
      ls = lsin
      t_tens = 1
      alpha1 = alpha1_
      dx1 = dx1_
      rho_f1 = rho1; rho_f2 = rho2
      mu_f1 = mu1; mu_f2 = mu2
      call heaviside()
      t_tens = 0
      hsout = hs
      rhoout = rho
      muout = ls_mu
      hsxout = hsx
      hsyout = hsy
      hszout = hsz
   end subroutine ref_heaviside_tens
#endif

end module suli_ref_bridge
