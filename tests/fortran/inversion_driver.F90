!##############################################################################
! Test driver for the Fortran source-inversion entry points (Ash3d_Adjoint):
! a forward run through the drop-in modules with the state resident on the
! GPU, source insertion at the vent every step and column loads at the
! observed steps, then the reverse sweep of the discrete adjoint.
!
! usage: inversion_driver in.bin out.bin
! in.bin (stream): the shim_driver case (its nsteps is ignored), then int32
!   ivent, jvent, n_heights, n_periods, steps_per_period, obs_every;
!   float64 frac(ns), dz(nz), s(n_periods, n_heights), w(nx, ny, n_obs).
! out.bin: float64 J, loads(nx, ny, n_obs), grad(n_periods, n_heights).
!##############################################################################

      program inversion_driver

      use precis_param
      use global_param,      only : useDiffusion
      use mesh
      use solution
      use time_data
      use Topography
      use Diffusion,         only : kx, ky, kz, Imp_fac, DiffuseHorz, DiffuseVert
      use AdvectionHorz,     only : AdvectHorz
      use AdvectionVert_DCU, only : AdvectVert
      use ash3d_tr_c,        only : ash3d_tr_resident
      use Ash3d_Adjoint

      implicit none

      interface
        subroutine Set_BC(bc_code)
          integer,intent(in) :: bc_code
        end subroutine Set_BC
        subroutine ash3d_tr_glue()
        end subroutine ash3d_tr_glue
        subroutine ash3d_tr_source_add(ivent, jvent, flux)
          integer, intent(in) :: ivent, jvent
          real(kind=8), intent(in) :: flux(*)
        end subroutine ash3d_tr_source_add
        subroutine ash3d_tr_detach()
        end subroutine ash3d_tr_detach
      end interface

      character(len=512) :: fin, fout
      integer :: hdr(12), ihdr(6), nsteps, itime, n, h, p, iobs, nobs
      integer :: nh, np, spp, obs_every
      real(kind=8) :: rhdr(2), J
      integer, allocatable :: aloft(:)
      real(kind=8), allocatable :: frac(:), dz(:), s(:,:), w(:,:,:), loads(:,:,:), G(:,:,:), grad(:,:), flux(:,:)

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(10, file=trim(fin), access='stream', form='unformatted', status='old')
      read(10) hdr
      nxmax = hdr(1); nymax = hdr(2); nzmax = hdr(3); nsmax = hdr(4)
      IsPeriodic = hdr(5) /= 0
      ZScaling_ID = hdr(6)
      useDiffusion = hdr(7) /= 0
      read(10) rhdr
      dt = rhdr(1)
      dt_ip = real(dt, ip)
      Imp_fac = rhdr(2)
      allocate(aloft(nsmax), IsAloft(nsmax))
      read(10) aloft
      IsAloft = aloft /= 0

      allocate(concen_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2,1:nsmax,0:1))
      allocate(vx_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2), vy_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2))
      allocate(vz_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2), vh_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2))
      allocate(vf_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2,1:nsmax))
      allocate(kappa_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2), sigma_nx_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2))
      allocate(sigma_ny_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2), sigma_nz_pd(-1:nxmax+2,-1:nymax+2,-1:nzmax+2))
      allocate(kx(0:nxmax+1,0:nymax+1,0:nzmax+1), ky(0:nxmax+1,0:nymax+1,0:nzmax+1))
      allocate(kz(0:nxmax+1,0:nymax+1,0:nzmax+1))
      allocate(DelDxonD_cc(-1:nxmax+2,-1:nymax+2), DelDyonD_cc(-1:nxmax+2,-1:nymax+2))
      allocate(outflow_xz1_pd(-1:nxmax+2,-1:nzmax+2,1:nsmax), outflow_xz2_pd(-1:nxmax+2,-1:nzmax+2,1:nsmax))
      allocate(outflow_yz1_pd(-1:nymax+2,-1:nzmax+2,1:nsmax), outflow_yz2_pd(-1:nymax+2,-1:nzmax+2,1:nsmax))
      allocate(outflow_xy1_pd(-1:nxmax+2,-1:nymax+2,1:nsmax), outflow_xy2_pd(-1:nxmax+2,-1:nymax+2,1:nsmax))
      allocate(DepositGranularity(nxmax,nymax,nsmax))
      read(10) concen_pd, vx_pd, vy_pd, vz_pd, vh_pd, vf_pd, kappa_pd, sigma_nx_pd, sigma_ny_pd, &
               sigma_nz_pd, kx, ky, kz, DelDxonD_cc, DelDyonD_cc, outflow_xz1_pd, outflow_xz2_pd, &
               outflow_yz1_pd, outflow_yz2_pd, outflow_xy1_pd, outflow_xy2_pd, DepositGranularity
      read(10) ihdr
      ivent = ihdr(1); jvent = ihdr(2); nh = ihdr(3); np = ihdr(4); spp = ihdr(5); obs_every = ihdr(6)
      nsteps = np * spp
      nobs = nsteps / obs_every
      allocate(frac(nsmax), dz(nzmax), s(np, nh), w(nxmax, nymax, nobs), loads(nxmax, nymax, nobs))
      allocate(G(nzmax+1, nsmax, np), grad(np, nh), flux(nzmax+1, nsmax))
      read(10) frac, dz, s, w
      close(10)
      allocate(dz_vec_pd(-1:nzmax+2))
      dz_vec_pd = 0.0_ip
      dz_vec_pd(1:nzmax) = dz

      ! forward run, every step input kept
      ash3d_tr_resident = .true.
      call Adjoint_Checkpoints(nsteps)
      call Adjoint_Init(np)
      J = 0.0d0
      iobs = 0
      do itime = 1, nsteps
        p = (itime - 1) / spp + 1
        flux = 0.0d0
        do n = 1, nsmax
          flux(1:nh, n) = s(p, 1:nh) * frac(n)
        enddo
        call ash3d_tr_glue()
        call ash3d_tr_source_add(ivent, jvent, flux)
        call Adjoint_Checkpoint(itime, CKPT_SAVE)
        call Set_BC(1)
        call AdvectHorz(itime)
        call AdvectVert
        if (useDiffusion) then
          call Set_BC(2)
          call DiffuseHorz(itime)
          call DiffuseVert
        endif
        if (mod(itime, obs_every) == 0) then
          iobs = iobs + 1
          call Column_Load(loads(:,:,iobs))
          J = J + sum(w(:,:,iobs) * loads(:,:,iobs))
        endif
      enddo

      ! reverse sweep
      call Adjoint_Init(np)
      do itime = nsteps, 1, -1
        if (mod(itime, obs_every) == 0) call Adjoint_Observe_Load(w(:,:,itime / obs_every))
        call Adjoint_Checkpoint(itime, CKPT_LOAD)
        call Adjoint_Step(itime)
        call Adjoint_Source(ivent, jvent, (itime - 1) / spp + 1)
      enddo
      call Adjoint_Source_Gradient(G)
      grad = 0.0d0
      do p = 1, np
        do h = 1, nh
          do n = 1, nsmax
            grad(p, h) = grad(p, h) + frac(n) * G(h, n, p)
          enddo
        enddo
      enddo
      call ash3d_tr_detach()

      open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(11) J, loads, grad
      close(11)
      end program inversion_driver
