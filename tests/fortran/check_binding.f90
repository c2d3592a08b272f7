! Links a Fortran program against libtandem_hip.so through the tandem_hip
! module and makes calls that need no GPU: without a device tp_create must
! fail with a message (the library never falls back to a CPU path), and
! calls on a null context must be rejected with TP_E_ARG.
program check_binding
  use iso_c_binding, only: c_ptr, c_null_ptr, c_int32_t, c_int64_t, c_double
  use tandem_hip
  implicit none
  type(c_ptr) :: ctx
  integer(c_int32_t) :: rc
  integer(c_int32_t) :: chunks(6)
  real(c_double) :: s(16)

  rc = tp_create(ctx)
  if (rc == TP_OK) then
    write(*, '(a)') 'tp_create: device present'
    call tp_check(tp_destroy(ctx), 'tp_destroy')
  else
    write(*, '(a,i0,a,a)') 'tp_create failed as expected, rc=', rc, ': ', tp_error_message()
  end if
  rc = tp_deriv(c_null_ptr, 1_c_int32_t, 1_c_int32_t, 1_c_int32_t)
  if (rc /= TP_E_ARG) error stop 2
  rc = tp_get_scalars(c_null_ptr, 0_c_int32_t, s)
  if (rc /= TP_E_ARG) error stop 3
  rc = tp_col_config(c_null_ptr, chunks)
  if (rc /= TP_E_ARG) error stop 4
  write(*, '(a)') 'null context rejected: ' // tp_error_message()
  write(*, '(a)') 'binding ok'
end program check_binding
