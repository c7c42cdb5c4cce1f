!##############################################################################
! Test driver for the drop-in shims: fills the modules from a binary case
! file, runs the reference driver's hot-path sequence per time step
! (src/Ash3d.F90:376-378 glue, :491-510 calls) through the shim modules and
! writes the module arrays back.
!
! usage: shim_driver in.bin out.bin
! in.bin (stream): int32 nx,ny,nz,ns,periodic,zscaling,use_diffusion,nsteps,
!   resident,ivent,jvent,kvent; float64 dt, imp_fac; int32 aloft(ns); then concen_pd, vx_pd,
!   vy_pd, vz_pd, vh_pd, vf_pd, kappa_pd, sigma_nx/ny/nz_pd, kx, ky, kz,
!   DelDxonD_cc, DelDyonD_cc, outflow_xz1/2, outflow_yz1/2, outflow_xy1/2,
!   DepositGranularity.  out.bin: the same arrays after the run.
!##############################################################################

      program shim_driver

      use precis_param
      use mesh
      use solution
      use time_data
      use Topography
      use Diffusion,         only : kx, ky, kz, Imp_fac, DiffuseHorz, DiffuseVert
      use AdvectionHorz,     only : AdvectHorz
      use AdvectionVert_DCU, only : AdvectVert
      use ash3d_tr_c,        only : ash3d_tr_resident
#ifdef LINE_SOLVER_PARTITIONED
      use ash3d_tr_c,        only : ash3d_tr_ctx, ash3d_tr_check, c_ash3d_tr_tune, &
                                    ASH3D_TUNE_LINE_SOLVER, ASH3D_LINE_SOLVER_PARTITIONED
#endif

      implicit none

      interface
        subroutine Set_BC(bc_code)
          integer,intent(in) :: bc_code
        end subroutine Set_BC
        subroutine ash3d_tr_glue()
        end subroutine ash3d_tr_glue
        subroutine ash3d_tr_download(mask)
          integer, intent(in) :: mask
        end subroutine ash3d_tr_download
        subroutine ash3d_tr_detach()
        end subroutine ash3d_tr_detach
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      character(len=512) :: fin, fout
      integer :: hdr(12), nsteps, itime, n, use_diffusion, resident
      real(kind=8) :: rhdr(2)
      integer, allocatable :: aloft(:)

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(10, file=trim(fin), access='stream', form='unformatted', status='old')
      read(10) hdr
      nxmax = hdr(1); nymax = hdr(2); nzmax = hdr(3); nsmax = hdr(4)
      IsPeriodic = hdr(5) /= 0
      ZScaling_ID = hdr(6)
      use_diffusion = hdr(7)
      nsteps = hdr(8)
      resident = hdr(9)
      ivent = hdr(10); jvent = hdr(11); kvent = hdr(12)
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
      close(10)

      ash3d_tr_resident = resident /= 0
#ifdef LINE_SOLVER_PARTITIONED
      ! the one added call that selects the partitioned line solver
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_tune(ash3d_tr_ctx, ASH3D_TUNE_LINE_SOLVER, &
                                          ASH3D_LINE_SOLVER_PARTITIONED), 'tune')
#endif
      do itime = 1, nsteps
        if (ash3d_tr_resident) then
          call ash3d_tr_glue()
        else
          concen_pd(:,:,:,:,ts0) = concen_pd(:,:,:,:,ts1)
          concen_pd(:,:,:,1:nsmax,ts1) = 0.0_ip
        endif
        call Set_BC(1)
        call AdvectHorz(itime)
        call AdvectVert
        if (use_diffusion /= 0) then
          call Set_BC(2)
          call DiffuseHorz(itime)
          call DiffuseVert
        endif
      enddo
      if (ash3d_tr_resident) call ash3d_tr_download(255)
      call ash3d_tr_detach()

      open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(11) concen_pd, vx_pd, vy_pd, vz_pd, vh_pd, vf_pd, kappa_pd, sigma_nx_pd, sigma_ny_pd, &
                sigma_nz_pd, kx, ky, kz, DelDxonD_cc, DelDyonD_cc, outflow_xz1_pd, outflow_xz2_pd, &
                outflow_yz1_pd, outflow_yz2_pd, outflow_xy1_pd, outflow_xy2_pd, DepositGranularity
      close(11)
      n = 0
      end program shim_driver
