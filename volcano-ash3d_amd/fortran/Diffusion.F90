!##############################################################################
!
!  Diffusion module  (drop-in for src/Diffusion.F90)
!
!  Same public entities: diffusivity_horz, diffusivity_vert, Imp_fac,
!  Imp_DT_fac, kx, ky, kz, Allocate_Diff, Deallocate_Diff, DiffuseHorz,
!  DiffuseVert.  The sweeps (diffCN_* when useCN, diff_* otherwise) run on
!  the MI355X through ash3d_tr_diffuse_dir.
!
!##############################################################################

      module Diffusion

      use precis_param

      use iso_c_binding

      use ash3d_tr_c

      use global_param,  only : &
         useVarDiffH,useVarDiffV

      implicit none

      private

      public Allocate_Diff,             &
             Deallocate_Diff,           &
             DiffuseHorz,               &
             DiffuseVert

      real(kind=ip),public :: diffusivity_horz    ! horizontal diffusion coefficient (km2/hr)
      real(kind=ip),public :: diffusivity_vert    ! vertical diffusion coefficient (km2/hr)

      ! 0.0 = Forward-Euler, 0.5 = Crank-Nicolson, 1.0 = Backward-Euler
      real(kind=ip),public  :: Imp_fac     = 0.5_ip
      real(kind=ip),public  :: Imp_DT_fac  = 4.0_ip

      real(kind=ip),dimension(:,:,:),allocatable,public :: kx
      real(kind=ip),dimension(:,:,:),allocatable,public :: ky
      real(kind=ip),dimension(:,:,:),allocatable,public :: kz

      interface
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      contains

      subroutine Allocate_Diff

      use mesh,          only : &
         nxmax,nymax,nzmax

      if(.not.allocated(kx))allocate(kx(0:nxmax+1,0:nymax+1,0:nzmax+1)); kx = diffusivity_horz
      if(.not.allocated(ky))allocate(ky(0:nxmax+1,0:nymax+1,0:nzmax+1)); ky = diffusivity_horz
      if(.not.allocated(kz))allocate(kz(0:nxmax+1,0:nymax+1,0:nzmax+1)); kz = diffusivity_vert

      end subroutine Allocate_Diff

      subroutine Deallocate_Diff

      if(allocated(kx)) deallocate(kx)
      if(allocated(ky)) deallocate(ky)
      if(allocated(kz)) deallocate(kz)

      end subroutine Deallocate_Diff

      subroutine diffuse_dir(dir, where)

      use global_param,  only : &
         useCN

      use mesh,          only : &
         nsmax

      use solution,      only : &
         IsAloft,imin,imax,jmin,jmax,kmin,kmax

      use time_data,     only : &
         dt_ip

      integer, intent(in) :: dir
      character(len=*), intent(in) :: where
      integer(c_int32_t) :: aloft(nsmax), bnd(6), cn

      call ash3d_tr_attach()
      call ash3d_tr_sync_in(where)
      aloft = merge(1_c_int32_t, 0_c_int32_t, IsAloft(1:nsmax))
      bnd = (/imin,imax,jmin,jmax,kmin,kmax/)
      cn = merge(1_c_int32_t, 0_c_int32_t, useCN)
      call ash3d_tr_check(c_ash3d_tr_diffuse_dir(ash3d_tr_ctx, int(dir,c_int32_t), cn, &
                          real(dt_ip,c_double), real(Imp_fac,c_double), aloft, bnd), where)
      call ash3d_tr_sync_out(where)

      end subroutine diffuse_dir

      subroutine DiffuseHorz(i)

      use mesh,          only : &
         nxmax,nymax,nzmax

      use solution,      only : &
         imin,imax,jmin,jmax,kmin,kmax

      integer :: i

#ifdef FAST_SUBGRID
      imin = max(1,imin-1)
      imax = min(nxmax,imax+1)
      jmin = max(1,jmin-1)
      jmax = min(nymax,jmax+1)
      kmin = max(1,kmin-1)
      kmax = min(nzmax,kmax+1)
#else
      imin = 1
      imax = nxmax
      jmin = 1
      jmax = nymax
      kmin = 1
      kmax = nzmax
#endif
      if(mod(i,2).eq.0) then
        call diffuse_dir(0, 'DiffuseHorz(x)')
        call diffuse_dir(1, 'DiffuseHorz(y)')
      else
        call diffuse_dir(1, 'DiffuseHorz(y)')
        call diffuse_dir(0, 'DiffuseHorz(x)')
      endif

      end subroutine DiffuseHorz

      subroutine DiffuseVert

      call diffuse_dir(2, 'DiffuseVert')

      end subroutine DiffuseVert

      end module Diffusion
