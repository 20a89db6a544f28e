!> ISO_C_BINDING interfaces to libmultall_hip (include/multall_hip.h).  This
!> is the binding a FORTRAN host adds to call the library in place of the
!> reference's `CALL LOOP` (multall-open-18.3.f:137).
module mh_hipif
  use iso_c_binding
  implicit none

  type, bind(C) :: mh_dims
    integer(c_int) :: im, jm, km, nrows, nib1, njb1, nkb1, nib2, njb2, nkb2, nsblk
  end type mh_dims

  type, bind(C) :: mh_resid
    integer(c_int) :: nstep, ifend
    real(c_float) :: emax, eavg, econt, vref, vmax, flow_in, flow_out
    integer(c_int) :: imax, jmax, kmax
  end type mh_resid

  type, bind(C) :: mh_perf
    integer(c_int) :: stage, j1, j2m1, jmix, compressor, pad_
    real(c_double) :: po_in, po_out, ps_in, ps_out, to_in, to_out, flow_in, flow_out, po_ratio
    real(c_double) :: eta_tt, eta_ts, eta_poly, reaction, reaction_hub, reaction_mid, reaction_tip, work_kw
  end type mh_perf

  interface
    integer(c_int) function mh_create(dims, ctx) bind(C, name='mh_create')
      import :: c_int, c_ptr, mh_dims
      type(mh_dims), intent(in) :: dims
      type(c_ptr), intent(out) :: ctx
    end function mh_create
    integer(c_int) function mh_destroy(ctx) bind(C, name='mh_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function mh_destroy
    type(c_ptr) function mh_last_error(ctx) bind(C, name='mh_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function mh_last_error
    integer(c_int) function mh_array_id(name) bind(C, name='mh_array_id')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: name(*)
    end function mh_array_id
    integer(c_int) function mh_scalar_id(name) bind(C, name='mh_scalar_id')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: name(*)
    end function mh_scalar_id
    integer(c_int) function mh_set_array(ctx, id, host, n) bind(C, name='mh_set_array')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id
      integer(c_long_long), value :: n
    end function mh_set_array
    integer(c_int) function mh_get_array(ctx, id, host, n) bind(C, name='mh_get_array')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id
      integer(c_long_long), value :: n
    end function mh_get_array
    integer(c_int) function mh_set_array_strided(ctx, id, host, ld_i, ld_j) bind(C, name='mh_set_array_strided')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id, ld_i, ld_j
    end function mh_set_array_strided
    integer(c_int) function mh_get_array_strided(ctx, id, host, ld_i, ld_j) bind(C, name='mh_get_array_strided')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id, ld_i, ld_j
    end function mh_get_array_strided
    integer(c_int) function mh_set_scalar(ctx, id, v) bind(C, name='mh_set_scalar')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), value :: v
    end function mh_set_scalar
    integer(c_int) function mh_loss_init(ctx) bind(C, name='mh_loss_init')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function mh_loss_init
    integer(c_int) function mh_set_xlength(ctx, theta, xllim_in, xllim_le, xllim_te, xllim_dn) &
        bind(C, name='mh_set_xlength')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: theta(*), xllim_in(*), xllim_le(*), xllim_te(*), xllim_dn(*)
    end function mh_set_xlength
    integer(c_int) function mh_start(ctx) bind(C, name='mh_start')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function mh_start
    integer(c_int) function mh_step(ctx, n, r) bind(C, name='mh_step')
      import :: c_int, c_ptr, mh_resid
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      type(mh_resid), intent(out) :: r
    end function mh_step
    integer(c_int) function mh_request_stop(ctx) bind(C, name='mh_request_stop')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function mh_request_stop
    integer(c_int) function mh_integrals(ctx, out, n) bind(C, name='mh_integrals')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), value :: n
    end function mh_integrals
    integer(c_int) function mh_mix_bconds(ctx, out, n) bind(C, name='mh_mix_bconds')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
      integer(c_long_long), value :: n
    end function mh_mix_bconds
    integer(c_int) function mh_eficool(ctx, out, cap, n) bind(C, name='mh_eficool')
      import :: c_int, c_ptr, mh_perf
      type(c_ptr), value :: ctx
      type(mh_perf), intent(out) :: out(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: n
    end function mh_eficool
  end interface

contains

  !> NUL-terminated copy of a FORTRAN string for the C side.
  function cname(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s)//c_null_char
  end function cname

end module mh_hipif
