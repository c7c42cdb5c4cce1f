!##############################################################################
!
! AdvectionHorz_DCU  (drop-in for src/AdvectionHorz_DCU.f90)
!
! Same module name and public routines (advect_x, advect_y, no arguments,
! reading the module state as the reference does); the sweeps run on the
! MI355X through the C ABI (ash3d_tr_advect_x / _y), each starting with
! Set_BC(1) like the reference (:115, :366).
!
!##############################################################################

      module AdvectionHorz_DCU

      use precis_param

      use iso_c_binding

      use ash3d_tr_c

      use mesh,          only : &
         nsmax

      use solution,      only : &
         IsAloft,imin,imax,jmin,jmax,kmin,kmax

      use time_data,     only : &
         dt

      implicit none

      private

      public advect_x,                  &
             advect_y

      interface
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      contains

      subroutine advect_x
      integer(c_int32_t) :: aloft(nsmax), bnd(6)
      call ash3d_tr_attach()
      call ash3d_tr_sync_in('advect_x')
      aloft = merge(1_c_int32_t, 0_c_int32_t, IsAloft(1:nsmax))
      bnd = (/imin,imax,jmin,jmax,kmin,kmax/)
      call ash3d_tr_check(c_ash3d_tr_advect_x(ash3d_tr_ctx, real(dt,c_double), aloft, bnd), 'advect_x')
      call ash3d_tr_sync_out('advect_x')
      end subroutine advect_x

      subroutine advect_y
      integer(c_int32_t) :: aloft(nsmax), bnd(6)
      call ash3d_tr_attach()
      call ash3d_tr_sync_in('advect_y')
      aloft = merge(1_c_int32_t, 0_c_int32_t, IsAloft(1:nsmax))
      bnd = (/imin,imax,jmin,jmax,kmin,kmax/)
      call ash3d_tr_check(c_ash3d_tr_advect_y(ash3d_tr_ctx, real(dt,c_double), aloft, bnd), 'advect_y')
      call ash3d_tr_sync_out('advect_y')
      end subroutine advect_y

      end module AdvectionHorz_DCU
