! mega_iface.f90 -- iso_c_binding interfaces of the C-ABI in include/mega.h,
! for a Fortran host that drives the explicit step the way MAIN drives
! HANDELER / CONSTRUCT_FINT (MAIN.F90:304-658).  All arrays keep the
! reference layouts: node vectors interleaved 3*(I-1)+K, tensors in Voigt
! order XX,YY,ZZ,YZ,XZ,XY, LOGICAL arrays as default-kind integers 0/1.
module mega_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MEGA_OK = 0, MEGA_ERR_RETURN_MAP = 1, MEGA_ERR_SINGULAR = 2, &
       MEGA_ERR_BINS = 3, MEGA_ERR_NOMEM = 4, MEGA_ERR_HIP = 5, MEGA_ERR_ARG = 6, MEGA_ERR_MATERIAL = 7
  integer(c_int32_t), parameter :: MEGA_FP64 = 0, MEGA_MIXED = 1
  integer(c_int32_t), parameter :: MEGA_F_X = 0, MEGA_F_V = 1, MEGA_F_A = 2, MEGA_F_D = 3, MEGA_F_DTOT = 4, &
       MEGA_F_UPHY = 5, MEGA_F_FINT = 6, MEGA_F_FEXT = 7, MEGA_F_STRESS = 8, MEGA_F_STRAIN = 9, &
       MEGA_F_EPS = 10, MEGA_F_STRAIN_EQ = 11, MEGA_F_WAVE_VEL = 12, MEGA_F_VPHY = 13, MEGA_F_DAMAGE_STATE = 14, &
       MEGA_F_APHY = 15, MEGA_F_VPHY_STEP = 16

  ! the CONTROL-module values the step reads (CONTROL.F90:27-95)
  type, bind(C) :: mega_control
     integer(c_int32_t) :: rk_degree, rk_cont, itype_int, ql
     real(c_double)     :: ql_coef, ql_len
     integer(c_int32_t) :: shsup, lagrangian, finite_strain, pdsearch
     real(c_double)     :: gravity(3), dist_max(3)
     integer(c_int32_t) :: precision, surrogate
     integer(c_int32_t) :: peridynamics, reserved
  end type mega_control

  interface
     integer(c_int) function mega_create(ctrl, n, x0, win, vol, mass, ebc, ebc_node, mat_type, prop, vinit, ctx) &
          bind(C, name="mega_create")
       import :: c_int, c_int32_t, c_double, c_ptr, mega_control
       type(mega_control), intent(in) :: ctrl
       integer(c_int32_t), value :: n
       real(c_double), intent(in) :: x0(*), win(*), vol(*), mass(*), prop(*), vinit(*)
       integer(c_int32_t), intent(in) :: ebc(*), ebc_node(*), mat_type(*)
       type(c_ptr), intent(out) :: ctx
     end function mega_create

     integer(c_int) function mega_destroy(ctx) bind(C, name="mega_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mega_destroy

     integer(c_int) function mega_search(ctx, x_cur, rebuild_bins) bind(C, name="mega_search")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: x_cur(*)
       integer(c_int32_t), value :: rebuild_bins
     end function mega_search

     integer(c_int) function mega_get_neighbors(ctx, gn, gstart, gstack, cap, nnz) bind(C, name="mega_get_neighbors")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: gn(*), gstart(*), gstack(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: nnz
     end function mega_get_neighbors

     integer(c_int) function mega_shapes(ctx, x_eval) bind(C, name="mega_shapes")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: x_eval(*)
     end function mega_shapes

     integer(c_int) function mega_interp(ctx, coef, phys) bind(C, name="mega_interp")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: coef(*)
       real(c_double), intent(out) :: phys(*)
     end function mega_interp

     integer(c_int) function mega_fint(ctx, linit, dinc, dinc_tot, x_cur, stress, strain, state, fint, fext, int_work) &
          bind(C, name="mega_fint")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: linit
       real(c_double), intent(in) :: dinc(*), dinc_tot(*), x_cur(*)
       real(c_double), intent(inout) :: stress(*), strain(*), state(*)
       real(c_double), intent(out) :: fint(*), fext(*)
       real(c_double), intent(inout) :: int_work
     end function mega_fint

     integer(c_int) function mega_step(ctx, dt, nsteps) bind(C, name="mega_step")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int32_t), value :: nsteps
     end function mega_step

     integer(c_int) function mega_step_auto(ctx, dlt_fac, nsteps, dlt) bind(C, name="mega_step_auto")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: dlt_fac
       integer(c_int32_t), value :: nsteps
       real(c_double), intent(out) :: dlt(*)
     end function mega_step_auto

     integer(c_int) function mega_download(ctx, field, host) bind(C, name="mega_download")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: field
       real(c_double), intent(out) :: host(*)
     end function mega_download

     integer(c_int) function mega_upload(ctx, field, host) bind(C, name="mega_upload")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: field
       real(c_double), intent(in) :: host(*)
     end function mega_upload

     integer(c_int) function mega_energies(ctx, e_int, e_kin, e_ext) bind(C, name="mega_energies")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: e_int, e_kin, e_ext
     end function mega_energies

     ! GET_ADIAL + GET_VOLUMES + GET_SMOOTHING_LENGTHS on the device (PRE_MODEL.F90:248-262);
     ! sm_len may be c_null_ptr
     integer(c_int) function mega_preprocess(nump, coo, numel, elcon, hex, win, vol, sm_len, dist_max) &
          bind(C, name="mega_preprocess")
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: nump, numel
       real(c_double), intent(in) :: coo(*)
       integer(c_int32_t), intent(in) :: elcon(*), hex(*)
       real(c_double), intent(out) :: win(*), vol(*)
       type(c_ptr), value :: sm_len
       real(c_double), intent(out) :: dist_max(3)
     end function mega_preprocess

     type(c_ptr) function mega_status_string(status) bind(C, name="mega_status_string")
       import :: c_int, c_ptr
       integer(c_int), value :: status
     end function mega_status_string
  end interface

contains

  ! status -> message, the way EXIT_PROGRAM / WARN report (EXIT_PROGRAM.F90)
  subroutine mega_check(status, what)
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (status == MEGA_OK) return
    call c_f_pointer(mega_status_string(status), msg, [256])
    write(*, '(A,A,I0,A)', advance='no') what, ': status ', status, ' ('
    do k = 1, 256
       if (msg(k) == c_null_char) exit
       write(*, '(A)', advance='no') msg(k)
    end do
    write(*, '(A)') ')'
    if (status /= MEGA_ERR_SINGULAR) stop 1
  end subroutine mega_check

end module mega_iface
