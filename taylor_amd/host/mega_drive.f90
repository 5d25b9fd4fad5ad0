! mega_drive.f90 -- Fortran host of the explicit step on the device: the
! time loop of MAIN (MAIN.F90:304-658) with HANDELER / CONSTRUCT_FINT /
! CONSTITUTION replaced by the C-ABI (mega_iface).  It takes the model
! arrays MAIN would hold after init (GCOO, GWIN, GVOL, LOCAL_MASS, LOCAL_EBC,
! GEBC_NODES, GMAT_TYPE, GPROP, LOCAL_VEL; MAIN.F90:130-302) either from the
! reference's own input (Control.dat and the LS-DYNA model file it names,
! module mega_deck) or from a binary stream file, advances the steps in
! chunks, and writes the final positions and plastic strain.
!
!   mega_drive --deck <Control.dat> <result.bin> [precision [nsteps [chunk]]]
!       precision 0 = FP64 (default), 1 = mixed; nsteps 0 or absent = the
!       reference's count (TIME += DLT until TIME > TIME_END, MAIN.F90:644-656).
!       With IAUTO_TS = 1 (peridynamic decks) each step's DLT comes from the
!       device (mega_step_auto) and TIME is summed here as MAIN sums it.
!       With 'unformatted' among the output variables, frames MEGAOUT_*.vtk
!       are written next to Control.dat on the reference's schedule: frame 0
!       before the loop, then after each step where the time since the last
!       frame exceeds TIME_OUTPUT (MAIN.F90:252-272,557-645).  At the end it
!       prints the length (z extent), the radius of the nodes fixed in z and
!       the volume-weighted L2 norm of the plastic strain.
!   mega_drive <model.bin> <result.bin>
!
! model.bin (stream, native endianness): int32 n, nsteps, chunk; real64 dt;
! mega_control; real64 X0(3n), win(3n), vol(n), mass(3n); int32 ebc(3n),
! ebc_node(n), mat_type(n); real64 prop(30n), vinit(3n).
program mega_drive
  use iso_c_binding
  use mega_iface
  use mega_deck
  use mega_vtk
  implicit none
  type(mega_control) :: ctrl
  type(deck_model) :: dm
  type(c_ptr) :: ctx
  integer(c_int32_t) :: n, nsteps, chunk, k, precision
  real(c_double) :: dt, e_int, e_kin, e_ext, t
  real(c_double), allocatable :: x0(:), win(:), vol(:), mass(:), prop(:), vinit(:), x(:), eps(:)
  integer(c_int32_t), allocatable :: ebc(:), ebc_node(:), mat_type(:)
  character(len=4096) :: fin, fout, arg
  integer :: done, ios, j, frame
  logical :: frames
  logical, allocatable :: outstep(:)
  real(c_double) :: tc, zlo, zhi, rad, num, den
  type(frame_fields) :: ff

  call get_command_argument(1, arg)
  if (trim(arg) == '--deck') then
     call get_command_argument(2, fin)
     call get_command_argument(3, fout)
     precision = MEGA_FP64; nsteps = 0; chunk = 500
     call get_command_argument(4, arg)
     if (len_trim(arg) > 0) read(arg, *, iostat=ios) precision
     call get_command_argument(5, arg)
     if (len_trim(arg) > 0) read(arg, *, iostat=ios) nsteps
     call get_command_argument(6, arg)
     if (len_trim(arg) > 0) read(arg, *, iostat=ios) chunk
     call deck_load(trim(fin), precision, dm)
     n = dm%n
     dt = dm%dt
     ctrl = dm%ctrl
     if (nsteps <= 0 .and. .not. dm%auto_ts) then
        t = 0.0d0
        do
           nsteps = nsteps + 1
           t = t + dt
           if (t > dm%time_end) exit
        end do
     end if
     write(*, '(A,I0,A,I0,A,ES14.6,A,I0)') 'deck: ', n, ' nodes, ', dm%numel, ' elements, dt ', dt, ', steps ', nsteps
     x0 = dm%x0; win = dm%win; vol = dm%vol; mass = dm%mass; ebc = dm%ebc; ebc_node = dm%ebc_node
     mat_type = dm%mat_type; prop = dm%prop; vinit = dm%vinit
  else
     fin = arg
     call get_command_argument(2, fout)
     open(10, file=trim(fin), access='stream', form='unformatted', status='old')
     read(10) n, nsteps, chunk, dt
     read(10) ctrl
     allocate(x0(3*n), win(3*n), vol(n), mass(3*n), ebc(3*n), ebc_node(n), mat_type(n), prop(30*n), vinit(3*n))
     read(10) x0, win, vol, mass, ebc, ebc_node, mat_type, prop, vinit
     close(10)
  end if

  call mega_check(mega_create(ctrl, n, x0, win, vol, mass, ebc, ebc_node, mat_type, prop, vinit, ctx), 'mega_create')
  frames = .false.
  if (allocated(dm%norm_win)) frames = dm%unf
  if (dm%auto_ts) then
     call auto_loop()
  else
     call fixed_loop()
  end if
  allocate(x(3*n), eps(n))
  call mega_check(mega_download(ctx, MEGA_F_X, x), 'mega_download')
  call mega_check(mega_download(ctx, MEGA_F_EPS, eps), 'mega_download')
  call mega_check(mega_destroy(ctx), 'mega_destroy')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) x, eps
  close(11)
  ! metrics: z extent, radius of the nodes fixed in z (the impact face), eqps L2
  zlo = x(3); zhi = x(3); rad = 0.0d0; num = 0.0d0; den = 0.0d0
  do j = 1, n
     zlo = min(zlo, x(3 * j)); zhi = max(zhi, x(3 * j))
     if (ebc(3 * j) == 1) rad = max(rad, sqrt(x(3 * j - 2) ** 2 + x(3 * j - 1) ** 2))
     num = num + vol(j) * eps(j) ** 2
     den = den + vol(j)
  end do
  write(*, '(A,ES24.16,A,ES24.16,A,ES24.16)') 'metrics: length ', zhi - zlo, ' radius ', rad, ' eqps_l2 ', &
       sqrt(num / den)

contains

  ! fixed DLT: the step count and the frame steps are known in advance
  subroutine fixed_loop()
    if (frames) then
       call vtk_notice_skipped(dm)
       allocate(outstep(nsteps))
       tc = 0.0d0
       do j = 1, nsteps          ! TIME_COUNTER: tested after the step, then advanced
          outstep(j) = tc > dm%time_output
          if (outstep(j)) tc = 0.0d0
          tc = tc + dt
       end do
       call get_frame(.true.)
       call vtk_frame(0, dm, ff)
       frame = 1
    end if
    done = 0
    do while (done < nsteps)
       k = min(chunk, nsteps - done)
       if (frames) then          ! stop at the next frame
          do j = done + 1, done + k
             if (outstep(j)) then
                k = j - done
                exit
             end if
          end do
       end if
       call mega_check(mega_step(ctx, dt, k), 'mega_step')
       done = done + k
       if (frames) then
          if (outstep(done)) then
             call get_frame(.false.)
             call vtk_frame(frame, dm, ff)
             frame = frame + 1
          end if
       end if
       call mega_check(mega_energies(ctx, e_int, e_kin, e_ext), 'mega_energies')
       write(*, '(A,I8,A,3ES14.6)') 'step ', done, '  energies (int, kin, ext)', e_int, e_kin, e_ext
    end do
  end subroutine fixed_loop

  ! IAUTO_TS = 1: each step's DLT comes back from the device; TIME and
  ! TIME_COUNTER are summed here in step order (MAIN.F90:641-642).  A frame
  ! follows a step when TIME_COUNTER, before that step's DLT is added, exceeds
  ! TIME_OUTPUT (MAIN.F90:557), so with frames the steps go one at a time;
  ! without them a call takes at most half the remaining time at the last DLT.
  subroutine auto_loop()
    real(c_double), allocatable :: dl(:)
    real(c_double) :: dlast
    integer(c_int32_t) :: kk
    integer :: jj
    logical :: out_now
    allocate(dl(max(chunk, 1)))
    if (frames) then
       call vtk_notice_skipped(dm)
       call get_frame(.true.)
       call vtk_frame(0, dm, ff)
       frame = 1
    end if
    t = 0.0d0; tc = 0.0d0; dlast = 0.0d0; done = 0; out_now = .false.
    do
       kk = 1
       if (.not. frames .and. dlast > 0.0d0) &
            kk = int(max(1.0d0, min(dble(max(chunk, 1)), 0.5d0 * (dm%time_end - t) / dlast)), c_int32_t)
       if (nsteps > 0) kk = min(kk, nsteps - done)
       call mega_check(mega_step_auto(ctx, dm%dt, kk, dl), 'mega_step_auto')
       do jj = 1, kk
          out_now = tc > dm%time_output
          if (out_now) tc = 0.0d0
          t = t + dl(jj)
          tc = tc + dl(jj)
       end do
       done = done + kk
       dlast = dl(kk)
       if (frames .and. out_now) then
          call get_frame(.false.)
          call vtk_frame(frame, dm, ff)
          frame = frame + 1
       end if
       if (nsteps > 0) then
          if (done >= nsteps) exit
       else if (t > dm%time_end) then
          exit
       end if
       if (.not. frames .or. mod(done, max(chunk, 1)) == 0) then
          call mega_check(mega_energies(ctx, e_int, e_kin, e_ext), 'mega_energies')
          write(*, '(A,I8,A,ES14.6,A,3ES14.6)') 'step ', done, '  DLT', dlast, '  energies (int, kin, ext)', &
               e_int, e_kin, e_ext
       end if
    end do
    call mega_check(mega_energies(ctx, e_int, e_kin, e_ext), 'mega_energies')
    write(*, '(A,I0,A,ES24.16,A,ES14.6,A,3ES14.6)') 'automatic time step: ', done, ' steps, TIME ', t, &
         ', last DLT', dlast, '; energies (int, kin, ext)', e_int, e_kin, e_ext
  end subroutine auto_loop

  ! the fields of one frame that the requested variables need; before the
  ! first step the reference writes the generalized displacement, velocity
  ! and acceleration (MAIN.F90:256-261)
  subroutine get_frame(initial)
    logical, intent(in) :: initial
    if (.not. allocated(ff%x)) then
       allocate(ff%x(3*n), ff%dsp(3*n), ff%vel(3*n), ff%acl(3*n), ff%fint(3*n), ff%stress(6*n), ff%strain(6*n), &
                ff%eps(n), ff%strain_eq(n), ff%wave(n), ff%dstate(4*n))
       ff%have_wave = ctrl%precision == MEGA_FP64 .or. ctrl%peridynamics == 1
       if (dm%outv(19) .and. .not. ff%have_wave) &
            write(*, '(A)') 'output variable max_wave_vel is not computed in mixed-precision RKPM; left out'
    end if
    call mega_check(mega_download(ctx, MEGA_F_X, ff%x), 'mega_download')
    if (initial) then
       if (dm%outv(2)) call mega_check(mega_download(ctx, MEGA_F_DTOT, ff%dsp), 'mega_download')
       if (dm%outv(3)) call mega_check(mega_download(ctx, MEGA_F_V, ff%vel), 'mega_download')
       if (dm%outv(4)) call mega_check(mega_download(ctx, MEGA_F_A, ff%acl), 'mega_download')
    else
       if (dm%outv(2)) call mega_check(mega_download(ctx, MEGA_F_UPHY, ff%dsp), 'mega_download')
       if (dm%outv(3)) call mega_check(mega_download(ctx, MEGA_F_VPHY_STEP, ff%vel), 'mega_download')
       if (dm%outv(4)) call mega_check(mega_download(ctx, MEGA_F_APHY, ff%acl), 'mega_download')
    end if
    if (dm%outv(5)) call mega_check(mega_download(ctx, MEGA_F_FINT, ff%fint), 'mega_download')
    if (dm%outv(21)) call mega_check(mega_download(ctx, MEGA_F_STRESS, ff%stress), 'mega_download')
    if (dm%outv(22)) then
       call mega_check(mega_download(ctx, MEGA_F_STRAIN, ff%strain), 'mega_download')
       call mega_check(mega_download(ctx, MEGA_F_STRAIN_EQ, ff%strain_eq), 'mega_download')
    end if
    if (dm%outv(23)) call mega_check(mega_download(ctx, MEGA_F_EPS, ff%eps), 'mega_download')
    if (dm%outv(19) .and. ff%have_wave) call mega_check(mega_download(ctx, MEGA_F_WAVE_VEL, ff%wave), 'mega_download')
    if (any(dm%outv(24:26))) call mega_check(mega_download(ctx, MEGA_F_DAMAGE_STATE, ff%dstate), 'mega_download')
  end subroutine get_frame

end program mega_drive
