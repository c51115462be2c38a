!> Fortran binding of libchapsim_hip (include/chapsim_hip.h) for the
!> reference solver.  The C ABI entry points are declared with bind(C);
!> the 18 Get_*_3D wrappers keep the reference's argument order
!> (src/basics_operations2.f90:270-312) with the domain argument replaced
!> by the library handle, and map a non-zero status to an error stop as
!> Print_error_msg does (src/tools_general.f90:10-20).
module chapsim_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CHS_OK = 0
  integer(c_int), parameter, public :: CHS_MEM_HOST = 0, CHS_MEM_DEVICE = 1
  integer(c_int), parameter, public :: CHS_PROJ_REFERENCE = 0, CHS_PROJ_EXACT = 1
  integer(c_int), parameter, public :: CHS_PREC_FP64 = 0, CHS_PREC_MIXED = 1
  integer(c_int), parameter, public :: CHS_POISSON_FISHPACK = 2
  integer(c_int), parameter, public :: CHS_PROPS_TABLE = 0, CHS_PROPS_MLP = 1
  ! operator kinds of chs_op3d
  integer(c_int), parameter, public :: CHS_MIDP_C2P = 0, CHS_MIDP_P2C = 1, CHS_1DER_C2C = 2, &
                                       CHS_1DER_P2P = 3, CHS_1DER_C2P = 4, CHS_1DER_P2C = 5

  !> chs_desc_t; uwall(component, side)
  type, bind(C), public :: chs_desc_t
    integer(c_int) :: nc(3)
    integer(c_int) :: periodic_y
    real(c_double) :: lxx, lyb, lyt, lzz
    integer(c_int) :: istret
    real(c_double) :: rstret
    integer(c_int) :: iacc
    real(c_double) :: dt
    real(c_double) :: ren, reninit
    integer(c_int) :: init_re_to
    integer(c_int) :: idriven
    integer(c_int) :: poisson
    real(c_double) :: uwall(3, 2)
    integer(c_int) :: precision
    integer(c_int) :: projection
  end type chs_desc_t

  !> chs_thermo_desc_t: the heated channel's property table and walls
  type, bind(C), public :: chs_thermo_desc_t
    type(c_ptr) :: props
    integer(c_int) :: mode
    real(c_double) :: t_wall(2)
    real(c_double) :: prandtl
    real(c_double) :: t_init
  end type chs_thermo_desc_t

  public :: chs_abi_version, chs_create, chs_destroy, chs_sync, chs_op3d, chs_poisson
  public :: chs_set_state, chs_get_state, chs_update_re, chs_substep, chs_run, chs_mass_check, chs_cfl, chs_bulk
  public :: chs_velocity_range
  public :: chs_divergence, chs_inspect_coeff, chs_inspect_op1d, chs_error_message, chs_check
  public :: chs_props_create, chs_props_destroy, chs_props_from_rhoh, chs_props_set_surrogate
  public :: chs_thermo_enable, chs_thermo_set_rhoh, chs_thermo_set_mode, chs_thermo_get, chs_energy_rhs
  public :: Get_x_midp_C2P_3D, Get_x_midp_P2C_3D, Get_x_1der_C2C_3D, Get_x_1der_P2P_3D, &
            Get_x_1der_C2P_3D, Get_x_1der_P2C_3D, Get_y_midp_C2P_3D, Get_y_midp_P2C_3D, &
            Get_y_1der_C2C_3D, Get_y_1der_P2P_3D, Get_y_1der_C2P_3D, Get_y_1der_P2C_3D, &
            Get_z_midp_C2P_3D, Get_z_midp_P2C_3D, Get_z_1der_C2C_3D, Get_z_1der_P2P_3D, &
            Get_z_1der_C2P_3D, Get_z_1der_P2C_3D

  interface
    integer(c_int) function chs_abi_version() bind(C, name="chs_abi_version")
      import :: c_int
    end function
    type(c_ptr) function chs_last_error() bind(C, name="chs_last_error")
      import :: c_ptr
    end function
    integer(c_int) function chs_create(desc, h) bind(C, name="chs_create")
      import :: c_int, c_ptr, chs_desc_t
      type(chs_desc_t), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function chs_destroy(h) bind(C, name="chs_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function chs_sync(h) bind(C, name="chs_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function chs_op3d(h, dir, op, iacc, ibc, fi, shp_in, fo, shp_out, fbc, mem) &
        bind(C, name="chs_op3d")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: dir, op, iacc, mem
      integer(c_int), intent(in) :: ibc(2), shp_in(3), shp_out(3)
      real(c_double), intent(in) :: fi(*)
      real(c_double), intent(inout) :: fo(*)
      type(c_ptr), value :: fbc
    end function
    integer(c_int) function chs_poisson(h, rhs, mem) bind(C, name="chs_poisson")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: rhs(*)
      integer(c_int), value :: mem
    end function
    integer(c_int) function chs_set_state(h, qx, qy, qz, pr, mem) bind(C, name="chs_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: qx(*), qy(*), qz(*), pr(*)
      integer(c_int), value :: mem
    end function
    integer(c_int) function chs_get_state(h, qx, qy, qz, pr, mem) bind(C, name="chs_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: qx(*), qy(*), qz(*), pr(*)
      integer(c_int), value :: mem
    end function
    integer(c_int) function chs_update_re(h, iter) bind(C, name="chs_update_re")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: iter
    end function
    integer(c_int) function chs_substep(h, isub) bind(C, name="chs_substep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: isub
    end function
    integer(c_int) function chs_run(h, iter0, nsteps) bind(C, name="chs_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: iter0, nsteps
    end function
    integer(c_int) function chs_mass_check(h, out) bind(C, name="chs_mass_check")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(3)
    end function
    ! Check_cfl_convection / Check_cfl_diffusion (src/tools_solver.f90:220-375)
    integer(c_int) function chs_cfl(h, out) bind(C, name="chs_cfl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(2)
    end function
    ! write_monitor_bulk values (src/io_monitor.f90:207-229)
    integer(c_int) function chs_bulk(h, out) bind(C, name="chs_bulk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(2)
    end function
    ! Find_max_min_3d of qx, qy, qz (src/tools_general.f90:1071-1113); CHS_ENAN on non-finite values
    integer(c_int) function chs_velocity_range(h, out) bind(C, name="chs_velocity_range")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(6)
    end function
    integer(c_int) function chs_divergence(h, div, mem) bind(C, name="chs_divergence")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: div(*)
      integer(c_int), value :: mem
    end function
    ! property table (buildup_property_relations_from_table, src/input_thermo.f90:617-693)
    integer(c_int) function chs_props_create(table, nrows, ref_t0, p) bind(C, name="chs_props_create")
      import :: c_int, c_ptr, c_double
      real(c_double), intent(in) :: table(*)
      integer(c_int), value :: nrows
      real(c_double), value :: ref_t0
      type(c_ptr), intent(out) :: p
    end function
    integer(c_int) function chs_props_destroy(p) bind(C, name="chs_props_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    ! ftp_refresh_thermal_properties_from_DH per element (src/input_thermo.f90:393-446);
    ! outputs are c_loc of arrays or c_null_ptr
    integer(c_int) function chs_props_from_rhoh(p, rhoh, n, h, d, m, k, t, b, cp, mode, mem) &
        bind(C, name="chs_props_from_rhoh")
      import :: c_int, c_long, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double), intent(in) :: rhoh(*)
      integer(c_long), value :: n
      type(c_ptr), value :: h, d, m, k, t, b, cp
      integer(c_int), value :: mode, mem
    end function
    integer(c_int) function chs_props_set_surrogate(p, mlp) bind(C, name="chs_props_set_surrogate")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, mlp
    end function
    ! variable-property heated channel (Solve_energy_eq + the thermo branches of Solve_momentum_eq)
    integer(c_int) function chs_thermo_enable(h, desc) bind(C, name="chs_thermo_enable")
      import :: c_int, c_ptr, chs_thermo_desc_t
      type(c_ptr), value :: h
      type(chs_thermo_desc_t), intent(in) :: desc
    end function
    integer(c_int) function chs_thermo_set_rhoh(h, rhoh, mem) bind(C, name="chs_thermo_set_rhoh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: rhoh(*)
      integer(c_int), value :: mem
    end function
    integer(c_int) function chs_thermo_set_mode(h, mode) bind(C, name="chs_thermo_set_mode")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function
    integer(c_int) function chs_thermo_get(h, rhoh, hh, t, k, d, m, gx, gy, gz, mem) bind(C, name="chs_thermo_get")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, rhoh, hh, t, k, d, m, gx, gy, gz
      integer(c_int), value :: mem
    end function
    integer(c_int) function chs_energy_rhs(h, isub, rhs, mem) bind(C, name="chs_energy_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: isub, mem
      real(c_double), intent(inout) :: rhs(*)
    end function
    real(c_double) function chs_inspect_coeff(fam, is_rhs, row, k, bc, acc) bind(C, name="chs_inspect_coeff")
      import :: c_int, c_double
      integer(c_int), value :: fam, is_rhs, row, k, bc, acc
    end function
    integer(c_int) function chs_inspect_op1d(desc, dir, op, iacc, ibc, fi, n_in, fo, n_out, fbc) &
        bind(C, name="chs_inspect_op1d")
      import :: c_int, c_double, c_ptr, chs_desc_t
      type(chs_desc_t), intent(in) :: desc
      integer(c_int), value :: dir, op, iacc, n_in, n_out
      integer(c_int), intent(in) :: ibc(2)
      real(c_double), intent(in) :: fi(*)
      real(c_double), intent(inout) :: fo(*)
      type(c_ptr), value :: fbc
    end function
  end interface

contains

  !> Last error message of the calling thread.
  function chs_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = chs_last_error()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  !> Stop with the library's message when status is not CHS_OK.
  subroutine chs_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= CHS_OK) then
      write (*, '(A)') 'chapsim_hip: '//where//': '//chs_error_message()
      error stop 1
    end if
  end subroutine

  subroutine op3d(h, dir, op, fi3d, fo3d, iacc, ibc, fbc2d)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: dir, op, iacc, ibc(2)
    real(c_double), intent(in), target, contiguous :: fi3d(:, :, :)
    real(c_double), intent(inout), target, contiguous :: fo3d(:, :, :)
    real(c_double), intent(in), target, contiguous, optional :: fbc2d(:, :, :)
    integer(c_int) :: si(3), so(3)
    type(c_ptr) :: pf
    si = int(shape(fi3d), c_int)
    so = int(shape(fo3d), c_int)
    pf = c_null_ptr
    if (present(fbc2d)) pf = c_loc(fbc2d)
    call chs_check(chs_op3d(h, dir, op, iacc, ibc, fi3d, si, fo3d, so, pf, CHS_MEM_HOST), 'chs_op3d')
  end subroutine

#define CHS_WRAP(NAME, DIR, OP) \
  subroutine NAME(fi3d, fo3d, h, iacc, ibc, fbc2d); \
    real(c_double), intent(in), contiguous :: fi3d(:, :, :); \
    real(c_double), intent(inout), contiguous :: fo3d(:, :, :); \
    type(c_ptr), intent(in) :: h; \
    integer, intent(in) :: iacc, ibc(2); \
    real(c_double), intent(in), contiguous, optional :: fbc2d(:, :, :); \
    call op3d(h, DIR, OP, fi3d, fo3d, int(iacc, c_int), int(ibc, c_int), fbc2d); \
  end subroutine

  CHS_WRAP(Get_x_midp_C2P_3D, 0_c_int, CHS_MIDP_C2P)
  CHS_WRAP(Get_x_midp_P2C_3D, 0_c_int, CHS_MIDP_P2C)
  CHS_WRAP(Get_x_1der_C2C_3D, 0_c_int, CHS_1DER_C2C)
  CHS_WRAP(Get_x_1der_P2P_3D, 0_c_int, CHS_1DER_P2P)
  CHS_WRAP(Get_x_1der_C2P_3D, 0_c_int, CHS_1DER_C2P)
  CHS_WRAP(Get_x_1der_P2C_3D, 0_c_int, CHS_1DER_P2C)
  CHS_WRAP(Get_y_midp_C2P_3D, 1_c_int, CHS_MIDP_C2P)
  CHS_WRAP(Get_y_midp_P2C_3D, 1_c_int, CHS_MIDP_P2C)
  CHS_WRAP(Get_y_1der_C2C_3D, 1_c_int, CHS_1DER_C2C)
  CHS_WRAP(Get_y_1der_P2P_3D, 1_c_int, CHS_1DER_P2P)
  CHS_WRAP(Get_y_1der_C2P_3D, 1_c_int, CHS_1DER_C2P)
  CHS_WRAP(Get_y_1der_P2C_3D, 1_c_int, CHS_1DER_P2C)
  CHS_WRAP(Get_z_midp_C2P_3D, 2_c_int, CHS_MIDP_C2P)
  CHS_WRAP(Get_z_midp_P2C_3D, 2_c_int, CHS_MIDP_P2C)
  CHS_WRAP(Get_z_1der_C2C_3D, 2_c_int, CHS_1DER_C2C)
  CHS_WRAP(Get_z_1der_P2P_3D, 2_c_int, CHS_1DER_P2P)
  CHS_WRAP(Get_z_1der_C2P_3D, 2_c_int, CHS_1DER_C2P)
  CHS_WRAP(Get_z_1der_P2C_3D, 2_c_int, CHS_1DER_P2C)

end module chapsim_hip
