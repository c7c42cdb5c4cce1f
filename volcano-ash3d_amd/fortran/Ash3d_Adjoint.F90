!##############################################################################
!
!  Ash3d_Adjoint
!
!  Fortran entry points of the discrete adjoint of the transport step, for a
!  source inversion driven from the reference's main program.  There is no
!  counterpart in the reference.  With the state resident on the GPU
!  (ash3d_tr_resident = .true.) a driver runs, for every step itime of the
!  forward pass,
!      call ash3d_tr_glue();  call ash3d_tr_source_add(ivent, jvent, flux)
!      call Adjoint_Checkpoint(slot, CKPT_SAVE)        ! the step's input
!      Set_BC(1); AdvectHorz(itime); AdvectVert; [Set_BC(2); DiffuseHorz(itime); DiffuseVert]
!      call Column_Load(load)                          ! observed steps
!  and then, for every step from the last one back to the first,
!      call Adjoint_Observe_Load(w)                    ! observed steps: lambda += L^T w
!      call Adjoint_Checkpoint(slot, CKPT_LOAD)
!      call Adjoint_Step(itime)                        ! lambda <- J(step input)^T lambda
!      call Adjoint_Source(ivent, jvent, period)       ! dJ/dflux of the step
!  and reads the source gradients with Adjoint_Source_Gradient.  Column
!  loads are ConcentrationCalculator's CloudLoad without its threshold
!  (src/Output_Vars.f90:1061-1066), over dz_vec_pd(1:nzmax).
!
!##############################################################################

      module Ash3d_Adjoint

      use precis_param

      use iso_c_binding

      use ash3d_tr_c

      implicit none

      private

      public Adjoint_Init,            &
             Adjoint_Checkpoints,     &
             Adjoint_Checkpoint,      &
             Adjoint_Step,            &
             Adjoint_Source,          &
             Adjoint_Source_Gradient, &
             Adjoint_Observe_Load,    &
             Column_Load

      integer, parameter, public :: CKPT_SAVE = 0
      integer, parameter, public :: CKPT_LOAD = 1

      interface
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      contains

      ! Allocates the adjoint buffers; lambda and n_periods source-gradient
      ! slots start at zero.
      subroutine Adjoint_Init(n_periods)
      integer, intent(in) :: n_periods
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_adj_init(ash3d_tr_ctx, int(n_periods, c_int32_t)), 'Adjoint_Init')
      end subroutine Adjoint_Init

      subroutine Adjoint_Checkpoints(nslots)
      integer, intent(in) :: nslots
      call ash3d_tr_attach()
      call ash3d_tr_check(c_ash3d_tr_checkpoints(ash3d_tr_ctx, int(nslots, c_int32_t)), 'Adjoint_Checkpoints')
      end subroutine Adjoint_Checkpoints

      ! Slots are numbered from 1.
      subroutine Adjoint_Checkpoint(slot, op)
      integer, intent(in) :: slot, op
      call ash3d_tr_check(c_ash3d_tr_checkpoint(ash3d_tr_ctx, int(slot - 1, c_int32_t), int(op, c_int32_t)), &
                          'Adjoint_Checkpoint')
      end subroutine Adjoint_Checkpoint

      subroutine Adjoint_Step(itime)

      use global_param,  only : &
         useDiffusion,useCN

      use mesh,          only : &
         nsmax

      use solution,      only : &
         IsAloft

      use time_data,     only : &
         dt_ip

      use Diffusion,     only : &
         Imp_fac

      integer, intent(in) :: itime
      integer(c_int32_t) :: aloft(nsmax)

      aloft = merge(1_c_int32_t, 0_c_int32_t, IsAloft(1:nsmax))
      call ash3d_tr_check(c_ash3d_tr_adj_step(ash3d_tr_ctx, int(itime, c_int32_t), real(dt_ip, c_double), &
                          merge(1_c_int32_t, 0_c_int32_t, useDiffusion), merge(1_c_int32_t, 0_c_int32_t, useCN), &
                          real(Imp_fac, c_double), aloft), 'Adjoint_Step')
      end subroutine Adjoint_Step

      ! Adjoint of ash3d_tr_source_add for this step, accumulated into
      ! source-gradient slot `period` (numbered from 1).
      subroutine Adjoint_Source(ivent, jvent, period)
      use time_data,     only : &
         dt
      integer, intent(in) :: ivent, jvent, period
      call ash3d_tr_check(c_ash3d_tr_adj_source(ash3d_tr_ctx, int(ivent, c_int32_t), int(jvent, c_int32_t), &
                          real(dt, c_double), int(period - 1, c_int32_t)), 'Adjoint_Source')
      end subroutine Adjoint_Source

      ! G(k, n, p) = sum over the steps of period p of dt*lambda(ivent,jvent,k,n)
      subroutine Adjoint_Source_Gradient(G)
      real(kind=8), intent(out) :: G(:,:,:)   ! (1:nzmax+1, 1:nsmax, 1:n_periods)
      call ash3d_tr_check(c_ash3d_tr_adj_source_read(ash3d_tr_ctx, G), 'Adjoint_Source_Gradient')
      end subroutine Adjoint_Source_Gradient

      subroutine Column_Load(load)
      use mesh,          only : &
         nzmax,dz_vec_pd
      real(kind=8), intent(out) :: load(:,:)   ! (1:nxmax, 1:nymax), t/km2
      real(kind=8) :: dz(nzmax)
      dz = dz_vec_pd(1:nzmax)
      call ash3d_tr_check(c_ash3d_tr_column_load(ash3d_tr_ctx, dz, load), 'Column_Load')
      end subroutine Column_Load

      subroutine Adjoint_Observe_Load(w)
      use mesh,          only : &
         nzmax,dz_vec_pd
      real(kind=8), intent(in) :: w(:,:)       ! dJ/dload, (1:nxmax, 1:nymax)
      real(kind=8) :: dz(nzmax)
      dz = dz_vec_pd(1:nzmax)
      call ash3d_tr_check(c_ash3d_tr_adj_column_load(ash3d_tr_ctx, dz, w), 'Adjoint_Observe_Load')
      end subroutine Adjoint_Observe_Load

      end module Ash3d_Adjoint
