! Host-only use of the Fortran binding: a coefficient-table entry and one
! planned y derivative through chs_inspect_op1d (no GPU needed).
program binding_check
  use iso_c_binding
  use chapsim_hip
  implicit none
  type(chs_desc_t) :: d
  integer(c_int) :: ibc(2), st, j
  real(c_double), target :: fi(24), fo(25), fbc(4)
  d%nc = [16_c_int, 24_c_int, 8_c_int]
  d%periodic_y = 0
  d%lxx = 1.0d0; d%lyb = -1.0d0; d%lyt = 1.0d0; d%lzz = 1.0d0
  d%istret = 2; d%rstret = 0.075d0
  d%iacc = 4; d%dt = 1.0d-3; d%ren = 1.0d0; d%reninit = 1.0d0
  d%init_re_to = 0; d%idriven = 0; d%poisson = CHS_POISSON_FISHPACK
  d%uwall = 0.0d0; d%precision = CHS_PREC_FP64; d%projection = CHS_PROJ_REFERENCE
  write (*, '(A,I0)') 'abi ', chs_abi_version()
  write (*, '(A,ES25.17)') 'coeff ', chs_inspect_coeff(2_c_int, 1_c_int, 3_c_int, 1_c_int, 0_c_int, 4_c_int)
  do j = 1, 24
    fi(j) = sin(0.3d0 * j) + 0.01d0 * j * j
  end do
  fbc = [0.5d0, -0.25d0, 0.0d0, 0.0d0]
  ibc = [5_c_int, 5_c_int]
  st = chs_inspect_op1d(d, 1_c_int, CHS_1DER_C2P, 4_c_int, ibc, fi, 24_c_int, fo, 25_c_int, c_loc(fbc))
  call chs_check(st, 'chs_inspect_op1d')
  do j = 1, 25
    write (*, '(ES25.17)') fo(j)
  end do
end program binding_check
