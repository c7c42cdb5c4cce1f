!##############################################################################
!
! AdvectionHorz  (drop-in for src/AdvectionHorz.f90)
!
! AdvectHorz(itoggle): the subgrid bounds (FAST_SUBGRID: get_minmax_index,
! src/Calc_Mesh.f90:432-542, evaluated by the library on the device copy of
! the concentrations, which is current in both residency modes; else the
! full grid), then the x/y sweeps in the order of the step parity (even: y
! then x; odd: x then y).
!
!##############################################################################

      module AdvectionHorz

      use precis_param

      implicit none

      private

      public AdvectHorz

      contains

      subroutine AdvectHorz(itoggle)

      use iso_c_binding

      use ash3d_tr_c

      use mesh,          only : &
         nxmax,nymax,nzmax

      use solution,      only : &
         imin,imax,jmin,jmax,kmin,kmax

      use AdvectionHorz_DCU, only : &
         advect_x,advect_y

#ifdef FAST_SUBGRID
      use mesh,          only : &
         ivent,jvent,kvent

      use Output_Vars,   only : &
         CLOUDCON_GRID_THRESH
#endif

      integer,intent(in) :: itoggle

#ifdef FAST_SUBGRID
      integer(c_int32_t) :: bnd(6)

      INTERFACE
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      END INTERFACE

      call ash3d_tr_attach()
      call ash3d_tr_sync_in('AdvectHorz')
      call ash3d_tr_check(c_ash3d_tr_minmax_index(ash3d_tr_ctx, int(ivent,c_int32_t), int(jvent,c_int32_t), &
                          int(kvent,c_int32_t), real(CLOUDCON_GRID_THRESH,c_double), bnd), 'get_minmax_index')
      imin = bnd(1)
      imax = bnd(2)
      jmin = bnd(3)
      jmax = bnd(4)
      kmin = bnd(5)
      kmax = bnd(6)
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
