! ISO_C_BINDING interfaces to include/fastmc_amd.h (the single-chain
! trial-move path).  Scalars by value, arrays by reference, opaque handle
! as type(c_ptr).
module fmc_abi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FMC_INSERT = 1, FMC_DELETE = 2, FMC_DISPLACE = 3

  interface
    integer(c_int) function fmc_ctx_create(ctx) bind(C, name='fmc_ctx_create')
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
    end function
    integer(c_int) function fmc_ctx_destroy(ctx) bind(C, name='fmc_ctx_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fmc_set_cell(ctx, imcon, cell) bind(C, name='fmc_set_cell')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: imcon
      real(c_double) :: cell(9)
    end function
    integer(c_int) function fmc_set_ewald(ctx, alpha, k1, k2, k3, epsq) &
        bind(C, name='fmc_set_ewald')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: alpha, epsq
      integer(c_int), value :: k1, k2, k3
    end function
    integer(c_int) function fmc_set_ewald_precision(ctx, eps, epsq) &
        bind(C, name='fmc_set_ewald_precision')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: eps, epsq
    end function
    integer(c_int) function fmc_set_cutoffs(ctx, rcut, delr, rvdw) &
        bind(C, name='fmc_set_cutoffs')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: rcut, delr, rvdw
    end function
    integer(c_int) function fmc_set_vdw(ctx, ntpatm, ntpvdw, a, b, keys, prm) &
        bind(C, name='fmc_set_vdw')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntpatm, ntpvdw
      integer(c_int) :: a(*), b(*), keys(*)
      real(c_double) :: prm(*)
    end function
    integer(c_int) function fmc_add_framework(ctx, n, x, y, z, q, t, frz) &
        bind(C, name='fmc_add_framework')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double) :: x(*), y(*), z(*), q(*)
      integer(c_int) :: t(*), frz(*)
    end function
    integer(c_int) function fmc_add_guest_type(ctx, ns, x, y, z, q, t, m, gt) &
        bind(C, name='fmc_add_guest_type')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ns
      real(c_double) :: x(*), y(*), z(*), q(*), m(*)
      integer(c_int) :: t(*)
      integer(c_int) :: gt
    end function
    integer(c_int) function fmc_add_guest_molecule(ctx, gt, x, y, z) &
        bind(C, name='fmc_add_guest_molecule')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: gt
      real(c_double) :: x(*), y(*), z(*)
    end function
    integer(c_int) function fmc_set_capacity(ctx, n) bind(C, name='fmc_set_capacity')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
    end function
    integer(c_int) function fmc_ctx_finalize(ctx) bind(C, name='fmc_ctx_finalize')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fmc_get_info(ctx, nk, volm, mxgrid, drewd, dlrpot) &
        bind(C, name='fmc_get_info')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int) :: nk, mxgrid
      real(c_double) :: volm, drewd, dlrpot
    end function
    integer(c_int) function fmc_guest_ewald3(ctx, gt, e3) bind(C, name='fmc_guest_ewald3')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: gt
      real(c_double) :: e3
    end function
    integer(c_int) function fmc_trial(ctx, kind, gt, imol, x, y, z, out) &
        bind(C, name='fmc_trial')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind, gt, imol
      real(c_double) :: x(*), y(*), z(*), out(3)
    end function
    integer(c_int) function fmc_ctx_timing(ctx, reset, out) bind(C, name='fmc_ctx_timing')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reset
      real(c_double) :: out(6)
    end function
    integer(c_int) function fmc_commit(ctx) bind(C, name='fmc_commit')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fmc_tail_delta(ctx, gt, sgn, elrc, delrc) &
        bind(C, name='fmc_tail_delta')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: gt, sgn
      real(c_double), value :: elrc
      real(c_double) :: delrc
    end function
    integer(c_int) function fmc_energy_full(ctx, out) bind(C, name='fmc_energy_full')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: out(6)
    end function
  end interface
end module fmc_abi
