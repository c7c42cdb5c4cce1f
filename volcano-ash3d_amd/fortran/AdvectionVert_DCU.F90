!##############################################################################
!
! AdvectionVert_DCU  (drop-in for src/AdvectionVert_DCU.f90)
!
! AdvectVert: the z sweep with settling (vz+vh+vf(n)), the deposit
! accumulation and DepositGranularity, on the MI355X.  advect_z and
! advect_sigma have the same body in the reference, so ZScaling_ID needs no
! dispatch here.
!
!##############################################################################

      module AdvectionVert_DCU

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

      public AdvectVert

      interface
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      contains

      subroutine AdvectVert
      integer(c_int32_t) :: aloft(nsmax), bnd(6)
      call ash3d_tr_attach()
      call ash3d_tr_sync_in('AdvectVert')
      aloft = merge(1_c_int32_t, 0_c_int32_t, IsAloft(1:nsmax))
      bnd = (/imin,imax,jmin,jmax,kmin,kmax/)
      call ash3d_tr_check(c_ash3d_tr_advect_z(ash3d_tr_ctx, real(dt,c_double), aloft, bnd), 'AdvectVert')
      call ash3d_tr_sync_out('AdvectVert')
      end subroutine AdvectVert

      end module AdvectionVert_DCU
