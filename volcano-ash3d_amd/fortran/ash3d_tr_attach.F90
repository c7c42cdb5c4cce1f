!##############################################################################
!
! ash3d_tr_attach
!
! Creates the device context from the reference's module arrays on the
! first call of any drop-in routine (after alloc_arrays and the mesh set-up,
! src/Ash3d.F90:230-240).  External on purpose: the drop-in Diffusion module
! owns kx, ky, kz (src/Diffusion.F90:62-64) and calls this routine, so a
! module here would be a dependency cycle.
!
!##############################################################################

      subroutine ash3d_tr_attach()

      use iso_c_binding

      use ash3d_tr_c

      use mesh,          only : &
         nxmax,nymax,nzmax,nsmax,IsPeriodic,ZScaling_ID,kappa_pd, &
         sigma_nx_pd,sigma_ny_pd,sigma_nz_pd

      use solution,      only : &
         concen_pd,vx_pd,vy_pd,vz_pd,vh_pd,vf_pd, &
         outflow_xz1_pd,outflow_xz2_pd,outflow_yz1_pd,outflow_yz2_pd, &
         outflow_xy1_pd,outflow_xy2_pd,DepositGranularity

      use Topography,    only : &
         DelDxonD_cc,DelDyonD_cc

      use Diffusion,     only : &
         kx,ky,kz

      implicit none

      type(ash3d_tr_desc) :: d

      if (c_associated(ash3d_tr_ctx)) return

      d%abi_version = ASH3D_TR_ABI_VERSION
      d%nx = nxmax
      d%ny = nymax
      d%nz = nzmax
      d%ns = nsmax
      d%is_periodic = merge(1, 0, IsPeriodic)
      d%zscaling_id = ZScaling_ID
      ! the limiter the reference selects at compile time (src/makefile:126)
#if defined(LIM_NONE)
      d%limiter = 0
#elif defined(LIM_LAXWEN)
      d%limiter = 1
#elif defined(LIM_BW)
      d%limiter = 2
#elif defined(LIM_FROMM)
      d%limiter = 3
#elif defined(LIM_MINMOD)
      d%limiter = 4
#elif defined(LIM_MC)
      d%limiter = 6
#else
      d%limiter = 5
#endif
      d%device = 0
      d%concen_pd   = ash3d_tr_addr(concen_pd)
      d%vx_pd       = ash3d_tr_addr(vx_pd)
      d%vy_pd       = ash3d_tr_addr(vy_pd)
      d%vz_pd       = ash3d_tr_addr(vz_pd)
      d%vh_pd       = ash3d_tr_addr(vh_pd)
      d%vf_pd       = ash3d_tr_addr(vf_pd)
      d%kappa_pd    = ash3d_tr_addr(kappa_pd)
      d%sigma_nx_pd = ash3d_tr_addr(sigma_nx_pd)
      d%sigma_ny_pd = ash3d_tr_addr(sigma_ny_pd)
      d%sigma_nz_pd = ash3d_tr_addr(sigma_nz_pd)
      d%kx = c_null_ptr
      d%ky = c_null_ptr
      d%kz = c_null_ptr
      if (allocated(kx)) d%kx = ash3d_tr_addr(kx)
      if (allocated(ky)) d%ky = ash3d_tr_addr(ky)
      if (allocated(kz)) d%kz = ash3d_tr_addr(kz)
      d%deldxond_cc = c_null_ptr
      d%deldyond_cc = c_null_ptr
      if (allocated(DelDxonD_cc)) d%deldxond_cc = ash3d_tr_addr(DelDxonD_cc)
      if (allocated(DelDyonD_cc)) d%deldyond_cc = ash3d_tr_addr(DelDyonD_cc)
      d%outflow_xz1_pd = ash3d_tr_addr(outflow_xz1_pd)
      d%outflow_xz2_pd = ash3d_tr_addr(outflow_xz2_pd)
      d%outflow_yz1_pd = ash3d_tr_addr(outflow_yz1_pd)
      d%outflow_yz2_pd = ash3d_tr_addr(outflow_yz2_pd)
      d%outflow_xy1_pd = ash3d_tr_addr(outflow_xy1_pd)
      d%outflow_xy2_pd = ash3d_tr_addr(outflow_xy2_pd)
      d%deposit_granularity = ash3d_tr_addr(DepositGranularity)

      call ash3d_tr_check(c_ash3d_tr_create(d, ash3d_tr_ctx), 'ash3d_tr_attach')

      end subroutine ash3d_tr_attach

!------------------------------------------------------------------------------
! Entry points for a driver that keeps the state on the GPU
! (ash3d_tr_resident = .true.): the step glue of src/Ash3d.F90:376-378, the
! vent-column source insertion of :428-431, and explicit syncs.
!------------------------------------------------------------------------------

      subroutine ash3d_tr_glue()
      use ash3d_tr_c
      implicit none
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_step_glue(ash3d_tr_ctx), 'ash3d_tr_glue')
      end subroutine ash3d_tr_glue

      subroutine ash3d_tr_source_add(ivent, jvent, flux)
      use iso_c_binding
      use ash3d_tr_c
      use time_data, only : dt
      implicit none
      integer, intent(in) :: ivent, jvent
      real(kind=8), intent(in) :: flux(*)   ! SourceNodeFlux(1:nzmax+1,1:nsmax)
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_source_column_add(ash3d_tr_ctx, int(ivent, c_int32_t), &
                          int(jvent, c_int32_t), real(dt, c_double), flux), 'ash3d_tr_source_add')
      end subroutine ash3d_tr_source_add

      subroutine ash3d_tr_download(mask)
      use iso_c_binding
      use ash3d_tr_c
      implicit none
      integer, intent(in) :: mask
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_sync_to_host(ash3d_tr_ctx, int(mask, c_int32_t)), 'ash3d_tr_download')
      end subroutine ash3d_tr_download

      subroutine ash3d_tr_upload(mask)
      use iso_c_binding
      use ash3d_tr_c
      implicit none
      integer, intent(in) :: mask
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_sync_from_host(ash3d_tr_ctx, int(mask, c_int32_t)), 'ash3d_tr_upload')
      end subroutine ash3d_tr_upload

      subroutine ash3d_tr_detach()
      use iso_c_binding
      use ash3d_tr_c
      implicit none
      integer(c_int) :: rc
      if (c_associated(ash3d_tr_ctx)) rc = c_ash3d_tr_destroy(ash3d_tr_ctx)
      ash3d_tr_ctx = c_null_ptr
      end subroutine ash3d_tr_detach
