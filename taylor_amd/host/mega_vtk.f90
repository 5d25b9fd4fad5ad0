! mega_vtk.f90 -- the reference program's 'unformatted' VTK frames
! (UNF_OUTPUT_STEP_VTK, VTK_OUTPUT_UNFORMATTED.F90:17-281, with the header,
! START_DATA and MID_DATA of VTK_OUTPUT.F90:331-538): the same records in the
! same order, written list-directed as the reference writes them, so a
! frame's text matches the reference's for equal values.  Kept as the
! reference has them: fixity is written when 'fint' is requested, and
! material_type is written twice.  Not computed on this path, and left out
! with a notice: characteristic_distance, pre_disp_force, IJKspace.
module mega_vtk
  use iso_c_binding
  use mega_deck
  implicit none
  private
  public :: frame_fields, vtk_frame, vtk_notice_skipped

  ! what the device hands over for one frame, node vectors interleaved
  type :: frame_fields
     real(c_double), allocatable :: x(:), dsp(:), vel(:), acl(:), fint(:), stress(:), strain(:), eps(:)
     real(c_double), allocatable :: strain_eq(:), wave(:), dstate(:)   ! dstate: EPS_SHEAR, EPS_VOL, D, DAM_SV
     logical :: have_wave = .false.   ! the wave speed is computed in FP64 mode only
  end type frame_fields

contains

  subroutine vtk_notice_skipped(m)
    type(deck_model), intent(in) :: m
    integer :: k
    do k = 1, NOUTV
       if (m%outv(k) .and. (k == 18 .or. k == 20 .or. k == 27)) &
            write(*, '(A,A,A)') 'output variable ', trim(out_names(k)), ' is not computed on this path; left out'
    end do
  end subroutine vtk_notice_skipped

  subroutine start_data(u)
    integer, intent(in) :: u
    write(u, *) ' '
  end subroutine start_data

  subroutine mid_data(u)
    integer, intent(in) :: u
    write(u, '(A20)') 'LOOKUP_TABLE default'
  end subroutine mid_data

  subroutine write_dbl(u, name, v, n, stride, off)
    integer, intent(in) :: u, n, stride, off
    character(*), intent(in) :: name
    real(c_double), intent(in) :: v(:)
    integer :: i
    call start_data(u)
    write(u, '(A8, A30, A10)') 'SCALARS ', name, 'double 1'
    call mid_data(u)
    do i = 1, n
       write(u, *) v(stride * (i - 1) + off)
    end do
  end subroutine write_dbl

  subroutine write_int(u, name, v, n, stride, off)
    integer, intent(in) :: u, n, stride, off
    character(*), intent(in) :: name
    integer(c_int32_t), intent(in) :: v(:)
    integer :: i, iv
    call start_data(u)
    write(u, '(A8, A30, A10)') 'SCALARS ', name, 'double 1'
    call mid_data(u)
    do i = 1, n
       iv = v(stride * (i - 1) + off)
       write(u, *) iv
    end do
  end subroutine write_int

  ! one frame MEGAOUT_<step>.vtk next to Control.dat
  subroutine vtk_frame(step, m, f)
    integer, intent(in) :: step
    type(deck_model), intent(in) :: m
    type(frame_fields), intent(in) :: f
    character(len=16) :: cfile
    integer :: u, i, n, ne
    integer(c_int32_t), allocatable :: ids(:)
    n = m%n
    ne = m%numel
    if (step < 10) then
       write(cfile, '(A11,I1,A4)') 'MEGAOUT_000', step, '.vtk'
    else if (step < 100) then
       write(cfile, '(A10,I2,A4)') 'MEGAOUT_00', step, '.vtk'
    else if (step < 1000) then
       write(cfile, '(A9,I3,A4)') 'MEGAOUT_0', step, '.vtk'
    else if (step < 10000) then
       write(cfile, '(A8,I4,A4)') 'MEGAOUT_', step, '.vtk'
    else
       call deck_fail('out of output file indices (10000 frames)')
    end if
    open(newunit=u, file=trim(m%dir)//cfile, status='replace')
    write(u, '(A26)') '# vtk DataFile Version 2.0 '
    write(u, '(A30)') '(MEGA VTK FILE)'
    write(u, '(A5)') 'ASCII'
    write(u, '(A25)') 'DATASET UNSTRUCTURED_GRID'
    write(u, '(A7,I10,A10)') 'POINTS ', n, 'double'
    do i = 1, n
       write(u, *) f%x(3 * i - 2), f%x(3 * i - 1), f%x(3 * i)
    end do
    write(u, *) ' '
    if (m%fem_output) then
       write(u, '(A6, I10, I20)') 'CELLS ', n + ne, n * 2 + ne * 9
    else
       write(u, '(A6, I10, I20)') 'CELLS ', n, n * 2
    end if
    do i = 1, n
       write(u, *) 1, i - 1
    end do
    if (m%fem_output) then
       do i = 1, ne
          write(u, *) 8, m%elcon(8 * i - 7:8 * i) - 1
       end do
    end if
    write(u, *) ' '
    if (m%fem_output) then
       write(u, '(A10, I8)') 'CELL_TYPES ', n + ne
    else
       write(u, '(A10, I8)') 'CELL_TYPES ', n
    end if
    do i = 1, n
       write(u, *) 1
    end do
    if (m%fem_output) then
       do i = 1, ne
          write(u, *) 12
       end do
    end if
    write(u, *) ' '
    write(u, *) ' '
    write(u, '(A11, I10)') 'POINT_DATA ', n
    allocate(ids(n))
    do i = 1, n
       ids(i) = i
    end do
    if (m%outv(8)) call write_int(u, 'material_type', m%mat_type, n, 1, 1)
    if (m%outv(1)) call write_int(u, 'node_id', ids, n, 1, 1)
    if (m%outv(7)) call write_int(u, 'bodyid', m%body_id, n, 1, 1)
    if (m%outv(8)) call write_int(u, 'material_type', m%mat_type, n, 1, 1)
    if (m%outv(5)) then
       call write_int(u, 'fixity_x', m%ebc, n, 3, 1)
       call write_int(u, 'fixity_y', m%ebc, n, 3, 2)
       call write_int(u, 'fixity_z', m%ebc, n, 3, 3)
    end if
    if (m%outv(2)) then
       call write_dbl(u, 'displacement_x', f%dsp, n, 3, 1)
       call write_dbl(u, 'displacement_y', f%dsp, n, 3, 2)
       call write_dbl(u, 'displacement_z', f%dsp, n, 3, 3)
    end if
    if (m%outv(3)) then
       call write_dbl(u, 'velocity_x', f%vel, n, 3, 1)
       call write_dbl(u, 'velocity_y', f%vel, n, 3, 2)
       call write_dbl(u, 'velocity_z', f%vel, n, 3, 3)
    end if
    if (m%outv(4)) then
       call write_dbl(u, 'acceleration_x', f%acl, n, 3, 1)
       call write_dbl(u, 'acceleration_y', f%acl, n, 3, 2)
       call write_dbl(u, 'acceleration_z', f%acl, n, 3, 3)
    end if
    if (m%outv(5)) then
       call write_dbl(u, 'fint_x', f%fint, n, 3, 1)
       call write_dbl(u, 'fint_y', f%fint, n, 3, 2)
       call write_dbl(u, 'fint_z', f%fint, n, 3, 3)
    end if
    if (m%outv(19) .and. f%have_wave) call write_dbl(u, 'max_wave_vel', f%wave, n, 1, 1)
    if (m%outv(24)) call write_dbl(u, 'eps_shear', f%dstate, n, 4, 1)
    if (m%outv(25)) call write_dbl(u, 'eps_vol', f%dstate, n, 4, 2)
    if (m%outv(23)) call write_dbl(u, 'eps', f%eps, n, 1, 1)
    if (m%outv(26)) call write_dbl(u, 'dam', f%dstate, n, 4, 3)
    if (m%outv(9)) then
       call write_dbl(u, 'initial_coordinate_x', m%x0, n, 3, 1)
       call write_dbl(u, 'initial_coordinate_y', m%x0, n, 3, 2)
       call write_dbl(u, 'initial_coordinate_z', m%x0, n, 3, 3)
    end if
    if (m%outv(10)) then
       call write_dbl(u, 'init_velocity_x', m%vinit_card, n, 3, 1)
       call write_dbl(u, 'init_velocity_y', m%vinit_card, n, 3, 2)
       call write_dbl(u, 'init_velocity_z', m%vinit_card, n, 3, 3)
    end if
    if (m%outv(11)) call write_dbl(u, 'normalized_window', m%norm_win, n, 1, 1)
    if (m%outv(17)) then
       call write_dbl(u, 'physical_window_x', m%win, n, 3, 1)
       call write_dbl(u, 'physical_window_y', m%win, n, 3, 2)
       call write_dbl(u, 'physical_window_z', m%win, n, 3, 3)
    end if
    if (m%outv(12)) call write_dbl(u, 'nodal_volume', m%vol, n, 1, 1)
    if (m%outv(21)) then
       call write_dbl(u, 'stress_xx', f%stress, n, 6, 1)
       call write_dbl(u, 'stress_yy', f%stress, n, 6, 2)
       call write_dbl(u, 'stress_zz', f%stress, n, 6, 3)
       call write_dbl(u, 'stress_yz', f%stress, n, 6, 4)
       call write_dbl(u, 'stress_xz', f%stress, n, 6, 5)
       call write_dbl(u, 'stress_xy', f%stress, n, 6, 6)
    end if
    if (m%outv(22)) then
       call write_dbl(u, 'strain_xx', f%strain, n, 6, 1)
       call write_dbl(u, 'strain_yy', f%strain, n, 6, 2)
       call write_dbl(u, 'strain_zz', f%strain, n, 6, 3)
       call write_dbl(u, 'strain_yz', f%strain, n, 6, 4)
       call write_dbl(u, 'strain_xz', f%strain, n, 6, 5)
       call write_dbl(u, 'strain_xy', f%strain, n, 6, 6)
       call write_dbl(u, 'strain_eq', f%strain_eq, n, 1, 1)
    end if
    if (m%outv(13)) call write_dbl(u, 'nodal_mass', m%mass, n, 3, 1)
    if (m%outv(14)) call write_dbl(u, 'Poissons_ratio', m%prop, n, 30, 1)
    if (m%outv(15)) call write_dbl(u, 'Youngs_modulus', m%prop, n, 30, 2)
    if (m%outv(16)) call write_dbl(u, 'density', m%prop, n, 30, 3)
    close(u)
  end subroutine vtk_frame

end module mega_vtk
