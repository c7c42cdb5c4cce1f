!##############################################################################
!
!  Set_BC(bc_code)  (drop-in for src/Set_BC.f90, an external subroutine)
!
!  bc_code 1: velocity ghost extrapolation and zero concentration ghosts
!  (periodic x copies); 2: Neumann mirror.  Any other code is fatal, as in
!  the reference (:119-124).
!
!##############################################################################

      subroutine Set_BC(bc_code)

      use iso_c_binding

      use ash3d_tr_c

      implicit none

      integer,intent(in) :: bc_code ! 1 for advection, 2 for diffusion

      interface
        subroutine ash3d_tr_attach()
        end subroutine ash3d_tr_attach
      end interface

      call ash3d_tr_attach()
      call ash3d_tr_sync_in('Set_BC')
      call ash3d_tr_check(c_ash3d_tr_set_bc(ash3d_tr_ctx, int(bc_code,c_int32_t)), 'Set_BC')
      call ash3d_tr_sync_out('Set_BC')

      end subroutine Set_BC
