!##############################################################################
!
! ash3d_tr_c
!
! ISO_C_BINDING interfaces to the C ABI of the MI355X transport step
! (include/ash3d_tr.h), the context handle shared by the drop-in modules,
! and the residency switch.
!
! With ash3d_tr_resident = .false. (default) every drop-in routine uploads
! the module arrays before its GPU work and downloads them after, so the
! unchanged reference driver (src/Ash3d.F90), which reads and writes
! concen_pd, the winds and the outflows on the host between calls, sees the
! same module state as with the Fortran routines.  A driver that calls
! ash3d_tr_glue / ash3d_tr_source_add and syncs explicitly before its own
! host reads sets ash3d_tr_resident = .true. and keeps the state on the GPU.
!
!##############################################################################

      module ash3d_tr_c

      use iso_c_binding

      implicit none

      private

      public ash3d_tr_desc, ash3d_tr_ctx, ash3d_tr_resident, ash3d_tr_check, &
             ash3d_tr_addr, ash3d_tr_sync_in, ash3d_tr_sync_out, &
             c_ash3d_tr_create, c_ash3d_tr_destroy, c_ash3d_tr_sync_from_host, &
             c_ash3d_tr_sync_to_host, c_ash3d_tr_set_bc, c_ash3d_tr_advect_x, &
             c_ash3d_tr_advect_y, c_ash3d_tr_advect_z, c_ash3d_tr_diffuse_dir, &
             c_ash3d_tr_step, c_ash3d_tr_step_glue, c_ash3d_tr_source_column_add, &
             c_ash3d_tr_last_error, ASH3D_TR_ABI_VERSION, ASH3D_SYNC_ALL, &
             c_ash3d_tr_checkpoints, c_ash3d_tr_checkpoint, c_ash3d_tr_adj_init, &
             c_ash3d_tr_adj_step, c_ash3d_tr_adj_source, c_ash3d_tr_adj_source_read, &
             c_ash3d_tr_column_load, c_ash3d_tr_adj_column_load, c_ash3d_tr_minmax_index, &
             c_ash3d_tr_met_push, c_ash3d_tr_met_interp, c_ash3d_tr_tune, &
             c_ash3d_tr_umbrella_source_add, c_ash3d_tr_umbrella_winds_add, &
             c_ash3d_tr_step_keep, c_ash3d_tr_adj_step_kept, &
             ASH3D_TUNE_LINE_SOLVER, ASH3D_LINE_SOLVER_EXACT, ASH3D_LINE_SOLVER_PARTITIONED

      integer(c_int32_t), parameter :: ASH3D_TR_ABI_VERSION = 1
      integer(c_int32_t), parameter :: ASH3D_SYNC_ALL = 255
      ! launch tuning (include/ash3d_tr.h): theta-scheme line solver
      integer(c_int32_t), parameter :: ASH3D_TUNE_LINE_SOLVER = 7
      integer(c_int32_t), parameter :: ASH3D_LINE_SOLVER_EXACT = 0
      integer(c_int32_t), parameter :: ASH3D_LINE_SOLVER_PARTITIONED = 1

      ! struct ash3d_tr_desc of include/ash3d_tr.h
      type, bind(C) :: ash3d_tr_desc
        integer(c_int32_t) :: abi_version
        integer(c_int32_t) :: nx, ny, nz, ns
        integer(c_int32_t) :: is_periodic
        integer(c_int32_t) :: zscaling_id
        integer(c_int32_t) :: limiter
        integer(c_int32_t) :: device
        type(c_ptr) :: concen_pd
        type(c_ptr) :: vx_pd, vy_pd, vz_pd, vh_pd, vf_pd
        type(c_ptr) :: kappa_pd, sigma_nx_pd, sigma_ny_pd, sigma_nz_pd
        type(c_ptr) :: kx, ky, kz
        type(c_ptr) :: deldxond_cc, deldyond_cc
        type(c_ptr) :: outflow_xz1_pd, outflow_xz2_pd
        type(c_ptr) :: outflow_yz1_pd, outflow_yz2_pd
        type(c_ptr) :: outflow_xy1_pd, outflow_xy2_pd
        type(c_ptr) :: deposit_granularity
      end type ash3d_tr_desc

      type(c_ptr), save :: ash3d_tr_ctx = c_null_ptr
      logical, save     :: ash3d_tr_resident = .false.

      interface
        integer(c_int) function c_ash3d_tr_create(desc, ctx) bind(C, name='ash3d_tr_create')
          import :: c_int, c_ptr, ash3d_tr_desc
          type(ash3d_tr_desc), intent(in) :: desc
          type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function c_ash3d_tr_destroy(ctx) bind(C, name='ash3d_tr_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_ash3d_tr_sync_from_host(ctx, mask) bind(C, name='ash3d_tr_sync_from_host')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: mask
        end function
        integer(c_int) function c_ash3d_tr_sync_to_host(ctx, mask) bind(C, name='ash3d_tr_sync_to_host')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: mask
        end function
        integer(c_int) function c_ash3d_tr_set_bc(ctx, code) bind(C, name='ash3d_tr_set_bc')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: code
        end function
        integer(c_int) function c_ash3d_tr_advect_x(ctx, dt, aloft, bnd) bind(C, name='ash3d_tr_advect_x')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: dt
          integer(c_int32_t), intent(in) :: aloft(*), bnd(6)
        end function
        integer(c_int) function c_ash3d_tr_advect_y(ctx, dt, aloft, bnd) bind(C, name='ash3d_tr_advect_y')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: dt
          integer(c_int32_t), intent(in) :: aloft(*), bnd(6)
        end function
        integer(c_int) function c_ash3d_tr_advect_z(ctx, dt, aloft, bnd) bind(C, name='ash3d_tr_advect_z')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: dt
          integer(c_int32_t), intent(in) :: aloft(*), bnd(6)
        end function
        integer(c_int) function c_ash3d_tr_diffuse_dir(ctx, dir, use_cn, dt, imp_fac, aloft, bnd) &
            bind(C, name='ash3d_tr_diffuse_dir')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: dir, use_cn
          real(c_double), value :: dt, imp_fac
          integer(c_int32_t), intent(in) :: aloft(*), bnd(6)
        end function
        integer(c_int) function c_ash3d_tr_step(ctx, itime, dt, use_diff, use_cn, imp_fac, aloft) &
            bind(C, name='ash3d_tr_step')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: itime, use_diff, use_cn
          real(c_double), value :: dt, imp_fac
          integer(c_int32_t), intent(in) :: aloft(*)
        end function
        integer(c_int) function c_ash3d_tr_step_glue(ctx) bind(C, name='ash3d_tr_step_glue')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_ash3d_tr_source_column_add(ctx, iv, jv, dt, flux) &
            bind(C, name='ash3d_tr_source_column_add')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: iv, jv
          real(c_double), value :: dt
          real(c_double), intent(in) :: flux(*)
        end function
        integer(c_int) function c_ash3d_tr_umbrella_source_add(ctx, iv, jv, dt, ibase, itop, ngs, stenc, flux) &
            bind(C, name='ash3d_tr_umbrella_source_add')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: iv, jv, ibase, itop, ngs
          real(c_double), value :: dt
          real(c_double), intent(in) :: stenc(*), flux(*)
        end function
        integer(c_int) function c_ash3d_tr_umbrella_winds_add(ctx, ibase, itop, uvx, uvy) &
            bind(C, name='ash3d_tr_umbrella_winds_add')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: ibase, itop
          real(c_double), intent(in) :: uvx(*), uvy(*)
        end function
        integer(c_int) function c_ash3d_tr_checkpoints(ctx, nslots) bind(C, name='ash3d_tr_checkpoints')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: nslots
        end function
        integer(c_int) function c_ash3d_tr_checkpoint(ctx, slot, op) bind(C, name='ash3d_tr_checkpoint')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: slot, op
        end function
        integer(c_int) function c_ash3d_tr_adj_init(ctx, nslots) bind(C, name='ash3d_tr_adj_init')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: nslots
        end function
        integer(c_int) function c_ash3d_tr_adj_step(ctx, itime, dt, use_diff, use_cn, imp_fac, aloft) &
            bind(C, name='ash3d_tr_adj_step')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: itime, use_diff, use_cn
          real(c_double), value :: dt, imp_fac
          integer(c_int32_t), intent(in) :: aloft(*)
        end function
        integer(c_int) function c_ash3d_tr_adj_step_kept(ctx, itime, dt, use_diff, use_cn, imp_fac, aloft, &
                                                         slot_b, slot_z) bind(C, name='ash3d_tr_adj_step_kept')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: itime, use_diff, use_cn, slot_b, slot_z
          real(c_double), value :: dt, imp_fac
          integer(c_int32_t), intent(in) :: aloft(*)
        end function
        integer(c_int) function c_ash3d_tr_step_keep(ctx, itime, dt, use_diff, use_cn, imp_fac, aloft, &
                                                     slot_b, slot_z) bind(C, name='ash3d_tr_step_keep')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: itime, use_diff, use_cn, slot_b, slot_z
          real(c_double), value :: dt, imp_fac
          integer(c_int32_t), intent(in) :: aloft(*)
        end function
        integer(c_int) function c_ash3d_tr_adj_source(ctx, iv, jv, dt, slot) bind(C, name='ash3d_tr_adj_source')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: iv, jv, slot
          real(c_double), value :: dt
        end function
        integer(c_int) function c_ash3d_tr_adj_source_read(ctx, out) bind(C, name='ash3d_tr_adj_source_read')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function c_ash3d_tr_column_load(ctx, dz, out) bind(C, name='ash3d_tr_column_load')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: dz(*)
          real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function c_ash3d_tr_adj_column_load(ctx, dz, w) bind(C, name='ash3d_tr_adj_column_load')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: dz(*), w(*)
        end function
        integer(c_int) function c_ash3d_tr_minmax_index(ctx, iv, jv, kv, thresh, bnd) &
            bind(C, name='ash3d_tr_minmax_index')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: iv, jv, kv
          real(c_double), value :: thresh
          integer(c_int32_t), intent(out) :: bnd(6)
        end function
        integer(c_int) function c_ash3d_tr_met_push(ctx, vx, vy, vz, vf) bind(C, name='ash3d_tr_met_push')
          import :: c_int, c_ptr, c_float
          type(c_ptr), value :: ctx
          real(c_float), intent(in) :: vx(*), vy(*), vz(*), vf(*)
        end function
        integer(c_int) function c_ash3d_tr_met_interp(ctx, frac, use_vf) bind(C, name='ash3d_tr_met_interp')
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: frac
          integer(c_int32_t), value :: use_vf
        end function
        integer(c_int) function c_ash3d_tr_tune(ctx, param, val) bind(C, name='ash3d_tr_tune')
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: param, val
        end function
        type(c_ptr) function c_ash3d_tr_last_error() bind(C, name='ash3d_tr_last_error')
          import :: c_ptr
        end function
      end interface

      contains

      ! Address of the first element of a contiguous array (module arrays of
      ! the reference are allocatables without TARGET; passing them to an
      ! assumed-size TARGET dummy gives the address of their storage).
      type(c_ptr) function ash3d_tr_addr(a)
        real(c_double), target, intent(in) :: a(*)
        ash3d_tr_addr = c_loc(a(1))
      end function ash3d_tr_addr

      ! Fatal error path of the reference (errlog + stop 1, e.g.
      ! src/Set_BC.f90:119-124), with the library's message.
      subroutine ash3d_tr_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (rc == 0) return
        text = ' '
        call c_f_pointer(c_ash3d_tr_last_error(), msg, [512])
        do i = 1, 512
          if (msg(i) == c_null_char) exit
          text(i:i) = msg(i)
        enddo
        write(0,*) 'ERROR in ', where, ' (ash3d_tr status ', rc, '): ', trim(text)
        stop 1
      end subroutine ash3d_tr_check

      subroutine ash3d_tr_sync_in(where)
        character(len=*), intent(in) :: where
        if (.not.ash3d_tr_resident) &
          call ash3d_tr_check(c_ash3d_tr_sync_from_host(ash3d_tr_ctx, ASH3D_SYNC_ALL), where)
      end subroutine ash3d_tr_sync_in

      subroutine ash3d_tr_sync_out(where)
        character(len=*), intent(in) :: where
        if (.not.ash3d_tr_resident) &
          call ash3d_tr_check(c_ash3d_tr_sync_to_host(ash3d_tr_ctx, ASH3D_SYNC_ALL), where)
      end subroutine ash3d_tr_sync_out

      end module ash3d_tr_c
