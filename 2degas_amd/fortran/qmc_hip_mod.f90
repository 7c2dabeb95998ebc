! qmc_hip_mod.f90 -- ISO_C_BINDING interface to lib2degas_hip.so.
!
! Mirrors include/qmc_hip.h one to one: bind(C) derived types for the
! qmc_tables / qmc_info / qmc_walkers structs and bind(C) interfaces for
! every entry point.  A Fortran host (the reference driver, or any other)
! uses this module to hand its deck tables to the device once and then to
! evaluate walkers through the C-ABI.
module qmc_hip_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QMC_OK = 0, QMC_ERR_ARG = 1, QMC_ERR_LIMIT = 2, &
       QMC_ERR_SINGULAR = 3, QMC_ERR_HIP = 4, QMC_ERR_NOTWIST = 5, QMC_ERR_STATE = 6
  integer(c_int), parameter :: QMC_PREC_FP64 = 0, QMC_PREC_MIXED = 1

  type, bind(C) :: qmc_tables
     integer(c_int32_t) :: ndim, ntypes
     integer(c_int32_t) :: np(2)
     real(c_double) :: hbs2m(2), el(2), drt
     integer(c_int32_t) :: ntab, ldut
     type(c_ptr) :: ut, ngrid, tail, iexp, nk_ewald
     integer(c_int32_t) :: iu2(4), iu3(4), iub(4), iv2(4)
     real(c_double) :: v2value(4), v0
     integer(c_int32_t) :: nk
     type(c_ptr) :: kvec
     integer(c_int32_t) :: irhok(2), nrhok, ntheta
     real(c_double) :: adrift, delta
  end type qmc_tables

  type, bind(C) :: qmc_info
     integer(c_int32_t) :: nptot, npnorm, n_scalar_props, n_props
     integer(c_int32_t) :: jetot, jltf, jacc, jpot0, jkin0, je2, jgg, jefp, jrhok
     integer(c_int32_t) :: jpot(2), jkin(2), jun(2)
     integer(c_int32_t) :: ng, has_det, reserved0
     integer(c_int64_t) :: dev_bytes, eval_bytes_per_walker, vmc_bytes_per_walker, &
          stage_bytes_per_walker
  end type qmc_info

  type, bind(C) :: qmc_walkers
     integer(c_int32_t) :: nw
     type(c_ptr) :: x, g, h, props, grad2_ph, age, rng, blk
  end type qmc_walkers

  type, bind(C) :: qmc_reptiles
     integer(c_int32_t) :: nr, ntau, nbuf
     type(c_ptr) :: x, g, props, age, w, ends, rage, pp_old, grad2_ph, rng, blk
  end type qmc_reptiles

  type, bind(C) :: qmc_dmc_params
     real(c_double) :: delta, etrial, mage
     integer(c_int32_t) :: mmstep, mmult, bm_order, reserved0
  end type qmc_dmc_params

  type, bind(C) :: qmc_dmc_state
     type(c_ptr) :: wate, wsum, src, rng2, counts
  end type qmc_dmc_state

  type, bind(C) :: qmc_sem_state
     type(c_ptr) :: ainv, dlogpsi, nacc
  end type qmc_sem_state

  type, bind(C) :: qmc_sem_job
     type(c_ptr) :: ctx, w, d
     integer(c_int32_t) :: j0, j1
     real(c_double) :: sigma
  end type qmc_sem_job

  type, bind(C) :: qmc_rmc_params
     real(c_double) :: delta, etrial, mage, mage_r
     integer(c_int32_t) :: ntauskip, mdelta, rejection, wpiu, gstorto, nodalaction, ecut, bm_order
     real(c_double) :: value_ecut
  end type qmc_rmc_params

  interface
     integer(c_int) function qmc_ctx_create(tables, device, ctx) bind(C, name='qmc_ctx_create')
       import :: c_int, c_ptr, qmc_tables
       type(qmc_tables), intent(in) :: tables
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
     end function qmc_ctx_create

     integer(c_int) function qmc_ctx_destroy(ctx) bind(C, name='qmc_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function qmc_ctx_destroy

     integer(c_int) function qmc_ctx_info(ctx, info) bind(C, name='qmc_ctx_info')
       import :: c_int, c_ptr, qmc_info
       type(c_ptr), value :: ctx
       type(qmc_info), intent(out) :: info
     end function qmc_ctx_info

     integer(c_int) function qmc_set_twist(ctx, gvec, ng) bind(C, name='qmc_set_twist')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: gvec(*)
       integer(c_int), value :: ng
     end function qmc_set_twist

     integer(c_int) function qmc_eval_batch_device_noalloc(ctx, nw, d_x, d_props, d_g, d_h, &
          d_grad2_ph, d_status, precision, stream) bind(C, name='qmc_eval_batch_device_noalloc')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, d_x, d_props, d_g, d_h, d_grad2_ph, d_status, stream
       integer(c_int), value :: nw, precision
     end function qmc_eval_batch_device_noalloc

     integer(c_int) function qmc_set_delta(ctx, delta) bind(C, name='qmc_set_delta')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: delta
     end function qmc_set_delta

     integer(c_int) function qmc_twist_gvec(kvec, nk, el, ntheta, ith, jth, gvec) &
          bind(C, name='qmc_twist_gvec')
       import :: c_int, c_double
       real(c_double), intent(in) :: kvec(*), el(*)
       integer(c_int), value :: nk, ntheta, ith, jth
       real(c_double), intent(out) :: gvec(*)
     end function qmc_twist_gvec

     integer(c_int) function qmc_eval_batch(ctx, nw, x, props, g, h, grad2_ph, status) &
          bind(C, name='qmc_eval_batch')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: nw
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(inout) :: props(*), g(*), h(*), grad2_ph(*)
       integer(c_int32_t), intent(out) :: status(*)
     end function qmc_eval_batch

     integer(c_int) function qmc_eval_batch_device(ctx, nw, x, props, g, h, grad2_ph, status, &
          precision, stream) bind(C, name='qmc_eval_batch_device')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, x, props, g, h, grad2_ph, status, stream
       integer(c_int), value :: nw, precision
     end function qmc_eval_batch_device

     integer(c_int) function qmc_vmc_start(ctx, w, stream) bind(C, name='qmc_vmc_start')
       import :: c_int, c_ptr, qmc_walkers
       type(c_ptr), value :: ctx, stream
       type(qmc_walkers), intent(in) :: w
     end function qmc_vmc_start

     integer(c_int) function qmc_vmc_sweep(ctx, w, nsteps, delta, mage, bm_order, trace_p, &
          trace_age, stream) bind(C, name='qmc_vmc_sweep')
       import :: c_int, c_ptr, c_double, qmc_walkers
       type(c_ptr), value :: ctx, trace_p, trace_age, stream
       type(qmc_walkers), intent(in) :: w
       integer(c_int), value :: nsteps, bm_order
       real(c_double), value :: delta, mage
     end function qmc_vmc_sweep

     integer(c_int) function qmc_vmc_status(ctx, nw, sticky) bind(C, name='qmc_vmc_status')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: nw
       integer(c_int32_t), intent(out) :: sticky(*)
     end function qmc_vmc_status

     integer(c_int) function qmc_dmc_sweep(ctx, w, d, nsteps, prm, stream) bind(C, name='qmc_dmc_sweep')
       import :: c_int, c_ptr, qmc_walkers, qmc_dmc_state, qmc_dmc_params
       type(c_ptr), value :: ctx, stream
       type(qmc_walkers), intent(in) :: w
       type(qmc_dmc_state), intent(in) :: d
       integer(c_int), value :: nsteps
       type(qmc_dmc_params), intent(in) :: prm
     end function qmc_dmc_sweep

     integer(c_int) function qmc_dmc_branch(ctx, nw, wate, rng2, src, counts, stream) &
          bind(C, name='qmc_dmc_branch')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, wate, rng2, src, counts, stream
       integer(c_int), value :: nw
     end function qmc_dmc_branch

     integer(c_int) function qmc_sem_refresh(ctx, w, d, precision, stream) &
          bind(C, name='qmc_sem_refresh')
       import :: c_int, c_ptr, qmc_walkers, qmc_sem_state
       type(c_ptr), value :: ctx, stream
       type(qmc_walkers), intent(in) :: w
       type(qmc_sem_state), intent(in) :: d
       integer(c_int), value :: precision
     end function qmc_sem_refresh

     integer(c_int) function qmc_sem_moves(ctx, w, d, j0, j1, sigma, precision, stream) &
          bind(C, name='qmc_sem_moves')
       import :: c_int, c_ptr, c_double, qmc_walkers, qmc_sem_state
       type(c_ptr), value :: ctx, stream
       type(qmc_walkers), intent(in) :: w
       type(qmc_sem_state), intent(in) :: d
       integer(c_int), value :: j0, j1, precision
       real(c_double), value :: sigma
     end function qmc_sem_moves

     integer(c_int) function qmc_sem_moves_packed(jobs, njobs, precision, stream) &
          bind(C, name='qmc_sem_moves_packed')
       import :: c_int, c_ptr, qmc_sem_job
       type(qmc_sem_job), intent(in) :: jobs(*)
       integer(c_int), value :: njobs, precision
       type(c_ptr), value :: stream
     end function qmc_sem_moves_packed

     integer(c_int) function qmc_autotune(ctx, nw, kchunk_out, ms_out, stream) bind(C, name='qmc_autotune')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, stream
       integer(c_int), value :: nw
       integer(c_int32_t), intent(out) :: kchunk_out(2)
       real(c_double), intent(out) :: ms_out(2)
     end function qmc_autotune

     integer(c_int) function qmc_sem_set_tile(ctx, precision, rows, threads_per_row) &
          bind(C, name='qmc_sem_set_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precision, rows, threads_per_row
     end function qmc_sem_set_tile

     integer(c_int) function qmc_rmc_start(ctx, r, prm, stream) bind(C, name='qmc_rmc_start')
       import :: c_int, c_ptr, qmc_reptiles, qmc_rmc_params
       type(c_ptr), value :: ctx, stream
       type(qmc_reptiles), intent(in) :: r
       type(qmc_rmc_params), intent(in) :: prm
     end function qmc_rmc_start

     integer(c_int) function qmc_rmc_sweep(ctx, r, prm, nsteps, trace, stream) &
          bind(C, name='qmc_rmc_sweep')
       import :: c_int, c_ptr, qmc_reptiles, qmc_rmc_params
       type(c_ptr), value :: ctx, trace, stream
       type(qmc_reptiles), intent(in) :: r
       type(qmc_rmc_params), intent(in) :: prm
       integer(c_int), value :: nsteps
     end function qmc_rmc_sweep

     integer(c_int) function qmc_timing(ctx, enable) bind(C, name='qmc_timing')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function qmc_timing

     integer(c_int) function qmc_timing_read(ctx, ms, nlaunch) bind(C, name='qmc_timing_read')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms(5)
       integer(c_int), intent(out) :: nlaunch
     end function qmc_timing_read

     type(c_ptr) function qmc_strerror(code) bind(C, name='qmc_strerror')
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function qmc_strerror

     type(c_ptr) function qmc_last_error() bind(C, name='qmc_last_error')
       import :: c_ptr
     end function qmc_last_error

     real(c_double) function qmc_flops_per_eval(nup, ndown, nk_ewald) &
          bind(C, name='qmc_flops_per_eval')
       import :: c_int, c_double
       integer(c_int), value :: nup, ndown, nk_ewald
     end function qmc_flops_per_eval
  end interface

contains

  ! Text of the library's last error as a Fortran string.
  function qmc_error_text() result(txt)
    character(len=256) :: txt
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    txt = ' '
    p = qmc_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do i = 1, 256
       if (cs(i) == c_null_char) exit
       txt(i:i) = cs(i)
    end do
  end function qmc_error_text

end module qmc_hip_mod
