!##############################################################################
! Test-only host modules for the drop-in shims.
!
! The shims are compiled inside the reference build against its own modules
! (src/Ash3d_VariableModules.f90, src/Topography.f90).  Those cannot be
! compiled here (they include a generated header and need MetReader), so
! this file declares the same module names with the same declarations for
! the entities the hot path uses (src/Ash3d_VariableModules.f90:23-37,
! :95-121, :271-273, :329-334, :600-620, :817-887, :1034-1072;
! src/Topography.f90:132-133), for the shim test driver only.
!##############################################################################

      module precis_param
      implicit none
      public
      integer, parameter,public :: ip = 8
      integer, parameter,public :: sp = selected_real_kind( 6,   37)
      integer, parameter,public :: dp = selected_real_kind(15,  307)
      end module precis_param

      module io_units
      implicit none
      integer :: io
      integer,dimension(2) :: VB = (/3,3/)
      integer,parameter :: verbosity_debug1 = 2
      integer,parameter :: verbosity_error  = 8
      end module io_units

      module global_param
      use precis_param
      implicit none
      real(kind=ip), parameter :: EPS_THRESH = 1.0e-10_ip
#ifdef CRANKNIC
      logical, parameter       :: useCN   = .true.
#else
      logical, parameter       :: useCN   = .false.
#endif
      logical                  :: useDiffusion     = .true.
      logical                  :: useVarDiffH      = .false.
      logical                  :: useVarDiffV      = .false.
      end module global_param

      module mesh
      use precis_param
      implicit none
      integer, parameter :: ts0 = 0
      integer, parameter :: ts1 = 1
      logical            :: IsPeriodic   = .false.
      integer            :: ZScaling_ID  = 0
      integer :: nxmax, nymax, nzmax, nsmax
      integer :: ivent, jvent, kvent
      real(kind=ip),dimension(:,:,:),allocatable :: kappa_pd
      real(kind=ip),dimension(:,:,:),allocatable :: sigma_nx_pd
      real(kind=ip),dimension(:,:,:),allocatable :: sigma_ny_pd
      real(kind=ip),dimension(:,:,:),allocatable :: sigma_nz_pd
      real(kind=ip),dimension(:)    ,allocatable :: dz_vec_pd
      end module mesh

      module solution
      use precis_param
      implicit none
      real(kind=ip),dimension(:,:,:)    ,allocatable :: vx_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: vy_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: vz_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: vh_pd
      real(kind=ip),dimension(:,:,:,:)  ,allocatable :: vf_pd
      real(kind=ip),dimension(:,:,:,:,:),allocatable :: concen_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_xz1_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_xz2_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_yz1_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_yz2_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_xy1_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: outflow_xy2_pd
      real(kind=ip),dimension(:,:,:)    ,allocatable :: DepositGranularity
      logical      ,dimension(:)        ,allocatable :: IsAloft
      integer :: imin,imax
      integer :: jmin,jmax
      integer :: kmin,kmax
      end module solution

      module Output_Vars
      use precis_param
      implicit none
      real(kind=ip), parameter :: CLOUDCON_GRID_THRESH = 1.0e-7_ip
      end module Output_Vars

      module time_data
      use precis_param
      implicit none
      real(kind=ip)      :: dt_ip
      real(kind=dp)      :: dt
      end module time_data

      module Topography
      use precis_param
      implicit none
      real(kind=ip),dimension(:,:)  ,allocatable,public :: DelDxonD_cc
      real(kind=ip),dimension(:,:)  ,allocatable,public :: DelDyonD_cc
      end module Topography
