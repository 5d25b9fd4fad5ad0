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
!       reference's count (TIME += DLT until TIME > TIME_END, MAIN.F90:644-656)
!   mega_drive <model.bin> <result.bin>
!
! model.bin (stream, native endianness): int32 n, nsteps, chunk; real64 dt;
! mega_control; real64 X0(3n), win(3n), vol(n), mass(3n); int32 ebc(3n),
! ebc_node(n), mat_type(n); real64 prop(30n), vinit(3n).
program mega_drive
  use iso_c_binding
  use mega_iface
  use mega_deck
  implicit none
  type(mega_control) :: ctrl
  type(deck_model) :: dm
  type(c_ptr) :: ctx
  integer(c_int32_t) :: n, nsteps, chunk, k, precision
  real(c_double) :: dt, e_int, e_kin, e_ext, t
  real(c_double), allocatable :: x0(:), win(:), vol(:), mass(:), prop(:), vinit(:), x(:), eps(:)
  integer(c_int32_t), allocatable :: ebc(:), ebc_node(:), mat_type(:)
  character(len=4096) :: fin, fout, arg
  integer :: done, ios

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
     if (nsteps <= 0) then
        t = 0.0d0
        do
           nsteps = nsteps + 1
           t = t + dt
           if (t > dm%time_end) exit
        end do
     end if
     write(*, '(A,I0,A,I0,A,ES14.6,A,I0)') 'deck: ', n, ' nodes, ', dm%numel, ' elements, dt ', dt, ', steps ', nsteps
     call move_alloc(dm%x0, x0); call move_alloc(dm%win, win); call move_alloc(dm%vol, vol)
     call move_alloc(dm%mass, mass); call move_alloc(dm%ebc, ebc); call move_alloc(dm%ebc_node, ebc_node)
     call move_alloc(dm%mat_type, mat_type); call move_alloc(dm%prop, prop); call move_alloc(dm%vinit, vinit)
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
  done = 0
  do while (done < nsteps)
     k = min(chunk, nsteps - done)
     call mega_check(mega_step(ctx, dt, k), 'mega_step')
     done = done + k
     call mega_check(mega_energies(ctx, e_int, e_kin, e_ext), 'mega_energies')
     write(*, '(A,I8,A,3ES14.6)') 'step ', done, '  energies (int, kin, ext)', e_int, e_kin, e_ext
  end do
  allocate(x(3*n), eps(n))
  call mega_check(mega_download(ctx, MEGA_F_X, x), 'mega_download')
  call mega_check(mega_download(ctx, MEGA_F_EPS, eps), 'mega_download')
  call mega_check(mega_destroy(ctx), 'mega_destroy')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) x, eps
  close(11)
end program mega_drive
