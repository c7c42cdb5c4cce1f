!##############################################################################
!
! AdvectionHorz  (drop-in for src/AdvectionHorz.f90)
!
! AdvectHorz(itoggle): the subgrid bounds (FAST_SUBGRID: the reference's own
! get_minmax_index, on host data; else the full grid), then the x/y sweeps
! in the order of the step parity (even: y then x; odd: x then y).
!
!##############################################################################

      module AdvectionHorz

      use precis_param

      implicit none

      private

      public AdvectHorz

      contains

      subroutine AdvectHorz(itoggle)

      use mesh,          only : &
         nxmax,nymax,nzmax

      use solution,      only : &
         imin,imax,jmin,jmax,kmin,kmax

      use AdvectionHorz_DCU, only : &
         advect_x,advect_y

      integer,intent(in) :: itoggle

#ifdef FAST_SUBGRID
      INTERFACE
        subroutine get_minmax_index
        end subroutine get_minmax_index
      END INTERFACE

      call get_minmax_index
#else
      imin = 1
      imax = nxmax
      jmin = 1
      jmax = nymax
      kmin = 1
      kmax = nzmax
#endif
      if(mod(itoggle,2).eq.0) then
        call advect_y
        call advect_x
      else
        call advect_x
        call advect_y
      endif

      end subroutine AdvectHorz

      end module AdvectionHorz
