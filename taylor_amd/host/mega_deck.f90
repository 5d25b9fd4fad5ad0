! mega_deck.f90 -- the reference program's own input on the Fortran host:
! the LS-DYNA model file and Control.dat, read with the reference's card
! grammar and assembled into the arrays mega_create takes, with the geometry
! preprocessing done on the device (mega_preprocess).
!
!   model file   *NODE (I8,3F16.0), *ELEMENT_SOLID (10I8), *SET_NODE_LIST
!                (set id, then 8I10 rows, entries up to the first 0)
!                                                   PRE_MODEL.F90:140-229
!   control      *MODEL, *TIME PARAMETERS, *QL PARAMETERS, *RK PARAMETERS,
!                *PERIDYNAMICS, *SUPPORT, *PERIODIC SEARCH, *FORMULATION
!                PARAMETERS, *INTEGRATION PARAMETERS, *GRAVITY PARAMETERS,
!                *OUTPUT PARAMETERS and their defaults
!                                                   PRE_CONTROL.F90:132-843
!   model cards  *SET NAME, *BOUNDARY CONDITION, *BODY DEFINITION,
!                *INITIAL CONDITION, *WINDOW DEFINITION, *MATERIALS
!                                                   READ.F90:96-658
! A card is its keyword line; after it come one skipped (comment) line and
! a value line, repeated as the card needs, read list-directed.  A node set
! is named by its number or by the name *SET NAME gave it (READ_SET_NO,
! READ.F90:662-735).  A value given twice to one node is an error
! (ASSIGN_TO_NODE.f90).  Then, as PRE_MODEL / PRE_CONTROL / STATE_INIT do:
! windows = ADIAL x normalised window, QL length = 0.1 x the smallest
! unscaled window, nodal mass = volume x density, no initial velocity on
! fixed dofs, EBC node = any fixity 1.
!
! Departures: elements after a block change are not preceded by a skipped
! line (the reference skips one line between blocks); a keyword line ends a
! node set's rows (the reference drops the line after a set unless it is
! another *SET_NODE_LIST); the block of an element is its part id, and a
! block is tetrahedral when its last element repeats id 7 as id 8, as
! PRE_MODEL.F90:231-239 decides.
module mega_deck
  use iso_c_binding
  use mega_iface
  implicit none
  private
  public :: deck_model, deck_load, deck_fail, NOUTV, out_names

  integer, parameter :: LL = 512, LLD = 4096

  ! *OUTPUT VARIABLES names (PRE_CONTROL.F90:139-183), in this order
  integer, parameter :: NOUTV = 27
  character(len=24), parameter :: out_names(NOUTV) = [character(len=24) :: &
       'node_id', 'displacement', 'velocity', 'acceleration', 'fint', 'fixity', 'bodyid', 'material_type', &
       'initial_coordinate', 'init_velocity', 'normalized_window', 'nodal_volume', 'nodal_mass', &
       'Poissons_ratio', 'Youngs_modulus', 'density', 'physical_window', 'characteristic_distance', &
       'max_wave_vel', 'pre_disp_force', 'stress', 'strain', 'eps', 'eps_shear', 'eps_vol', 'dam', 'IJKspace']

  type :: deck_model
     integer(c_int32_t) :: n = 0, numel = 0
     real(c_double) :: dt = 0.0d0, time_end = 0.0d0, time_output = 0.0d0
     logical :: auto_ts = .false.                     ! IAUTO_TS = 1: dt holds DLT_FAC
     type(mega_control) :: ctrl
     real(c_double), allocatable :: x0(:), win(:), vol(:), mass(:), prop(:), vinit(:), norm_win(:)
     real(c_double), allocatable :: vinit_card(:)     ! MODEL_VINIT as the cards give it (fixed dofs too)
     integer(c_int32_t), allocatable :: ebc(:), ebc_node(:), mat_type(:), body_id(:), elcon(:)
     logical :: unf = .false., fem_output = .false.   ! 'unformatted' frames; *FEM OUTPUT cells
     logical :: outv(NOUTV) = .false.                 ! requested output variables
     character(len=LLD) :: dir = ''                   ! directory of Control.dat (frames go there)
  end type deck_model


contains

  subroutine deck_fail(msg)
    character(*), intent(in) :: msg
    write(*, '(A,A)') 'deck error: ', msg
    stop 2
  end subroutine deck_fail

  ! ------------------------------------------------------------------ model file
  subroutine grow_d(a, need)
    real(c_double), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: need
    real(c_double), allocatable :: t(:)
    if (size(a) >= need) return
    allocate(t(max(need, 2 * size(a))))
    t = 0.0d0
    t(1:size(a)) = a
    call move_alloc(t, a)
  end subroutine grow_d

  subroutine grow_i(a, need)
    integer(c_int32_t), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: need
    integer(c_int32_t), allocatable :: t(:)
    if (size(a) >= need) return
    allocate(t(max(need, 2 * size(a))))
    t = 0
    t(1:size(a)) = a
    call move_alloc(t, a)
  end subroutine grow_i

  ! nodes coo(3n), elements elcon(8 numel) with part ids, node sets as
  ! set_id(ns), set_ptr(ns+1), set_node(:) (1-based positions)
  subroutine read_model_file(path, n, coo, numel, elcon, part, ns, set_id, set_ptr, set_node)
    character(*), intent(in) :: path
    integer, intent(out) :: n, numel, ns
    real(c_double), allocatable, intent(out) :: coo(:)
    integer(c_int32_t), allocatable, intent(out) :: elcon(:), part(:), set_id(:), set_ptr(:), set_node(:)
    character(len=LL) :: line
    integer :: u, ios, nid, ids(10), k, nse, sec
    real(c_double) :: xyz(3)
    logical, allocatable :: seen(:)
    integer, parameter :: S_NONE = 0, S_NODE = 1, S_ELEM = 2, S_SETHEAD = 3, S_SET = 4

    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call deck_fail('cannot open the model file '//trim(path))
    allocate(coo(3 * 4096), elcon(8 * 4096), part(4096), set_id(16), set_ptr(17), set_node(4096), seen(4096))
    coo = 0.0d0; seen = .false.
    n = 0; numel = 0; ns = 0; nse = 0; sec = S_NONE
    set_ptr(1) = 1
    do
       read(u, '(A)', iostat=ios) line
       if (ios /= 0) exit
       if (line(1:1) == '*') then
          select case (trim(line))
          case ('*NODE')
             sec = S_NODE
          case ('*ELEMENT_SOLID')
             sec = S_ELEM
          case ('*SET_NODE_LIST')
             sec = S_SETHEAD
          case default
             sec = S_NONE
          end select
          cycle
       end if
       if (line(1:1) == '$' .or. len_trim(line) == 0 .and. sec /= S_SET) cycle
       select case (sec)
       case (S_NODE)
          read(line, '(I8,3F16.0)', iostat=ios) nid, xyz
          if (ios /= 0 .or. nid < 1) call deck_fail('bad *NODE line: '//trim(line))
          call grow_d(coo, 3 * nid)
          if (size(seen) < nid) then
             block
               logical, allocatable :: t(:)
               allocate(t(max(nid, 2 * size(seen)))); t = .false.; t(1:size(seen)) = seen
               call move_alloc(t, seen)
             end block
          end if
          if (seen(nid)) call deck_fail('node defined twice')
          seen(nid) = .true.
          coo(3 * nid - 2:3 * nid) = xyz
          n = max(n, nid)
       case (S_ELEM)
          read(line, '(10I8)', iostat=ios) ids
          if (ios /= 0) call deck_fail('bad *ELEMENT_SOLID line: '//trim(line))
          numel = numel + 1
          call grow_i(elcon, 8 * numel)
          call grow_i(part, numel)
          elcon(8 * numel - 7:8 * numel) = ids(3:10)
          part(numel) = ids(2)
       case (S_SETHEAD)
          ns = ns + 1
          call grow_i(set_id, ns)
          call grow_i(set_ptr, ns + 1)
          read(line, *, iostat=ios) set_id(ns)
          if (ios /= 0) call deck_fail('bad *SET_NODE_LIST id line: '//trim(line))
          set_ptr(ns + 1) = set_ptr(ns)
          sec = S_SET
       case (S_SET)
          ids = 0
          read(line, '(8I10)', iostat=ios) ids(1:8)
          if (ios /= 0) call deck_fail('bad *SET_NODE_LIST row: '//trim(line))
          do k = 1, 8            ! entries up to the first 0 (PRE_MODEL.F90:205-213)
             if (ids(k) == 0) exit
             nse = nse + 1
             call grow_i(set_node, nse)
             set_node(nse) = ids(k)
          end do
          set_ptr(ns + 1) = nse + 1
       end select
    end do
    close(u)
    if (n == 0) call deck_fail('no *NODE section in '//trim(path))
    if (.not. all(seen(1:n))) call deck_fail('node ids are not 1..N')
  end subroutine read_model_file

  ! ------------------------------------------------------------------ control file
  subroutine read_lines(path, lines, nl)
    character(*), intent(in) :: path
    character(len=LL), allocatable, intent(out) :: lines(:)
    integer, intent(out) :: nl
    character(len=LL), allocatable :: t(:)
    character(len=LL) :: buf
    integer :: u, ios
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call deck_fail('cannot open '//trim(path))
    allocate(lines(256))
    nl = 0
    do
       read(u, '(A)', iostat=ios) buf
       if (ios /= 0) exit
       nl = nl + 1
       if (nl > size(lines)) then
          allocate(t(2 * size(lines))); t(1:size(lines)) = lines
          call move_alloc(t, lines)
       end if
       lines(nl) = buf
    end do
    close(u)
  end subroutine read_lines

  ! the value line `skip` lines below a card's keyword at line i (one comment
  ! line before each value line: value k sits at i + 2k)
  function val(lines, nl, i, k) result(s)
    character(len=LL), intent(in) :: lines(:)
    integer, intent(in) :: nl, i, k
    character(len=LL) :: s
    if (i + 2 * k > nl) call deck_fail('card '//trim(lines(i))//' is cut short')
    s = lines(i + 2 * k)
  end function val

  ! READ_SET_NO (READ.F90:662-735): a number is a *SET_NODE_LIST id; a name
  ! is looked up in the *SET NAME table and its set number then used as the
  ! position in the model file's list of sets, as the reference does
  integer function set_index(line, ns, set_id, names)
    character(*), intent(in) :: line
    integer, intent(in) :: ns
    integer(c_int32_t), intent(in) :: set_id(:)
    character(len=100), intent(in) :: names(:)
    integer :: ios, sno, i
    character(len=50) :: c
    c = line(1:50)
    read(c, '(I50)', iostat=ios) sno
    set_index = 0
    if (ios == 0) then
       do i = 1, ns
          if (set_id(i) == sno) set_index = i
       end do
       if (set_index == 0) call deck_fail('set number not found: '//trim(c))
    else
       do i = 1, size(names)
          if (names(i) == c) set_index = i
       end do
       if (set_index == 0 .or. set_index > ns) call deck_fail('set name not found: '//trim(c))
    end if
  end function set_index

  subroutine assign_int(existing, trial, what)
    integer(c_int32_t), intent(inout) :: existing
    integer, intent(in) :: trial
    character(*), intent(in) :: what
    if (existing /= -1) call deck_fail('multiple values of '//what//' assigned to a single node')
    existing = trial
  end subroutine assign_int

  subroutine assign_dbl(existing, trial, what)
    real(c_double), intent(inout) :: existing
    real(c_double), intent(in) :: trial
    character(*), intent(in) :: what
    if (abs(existing) > 1.0d-16) call deck_fail('multiple values of '//what//' assigned to a single node')
    existing = trial
  end subroutine assign_dbl

  ! props 25-27 = Q_PHI, K_PHI, Q_PSI of the Drucker-Prager cone from DP_TYPE,
  ! PHI (degrees), C = props 4-6 (READ.F90:420-483); Q_PSI is reset to 0
  subroutine drucker_prager_prep(props)
    real(c_double), intent(inout) :: props(30)
    integer :: dp_type
    real(c_double) :: phi, c, phi_rad, srt32, q_phi, k_phi
    dp_type = int(props(4))
    phi = props(5)
    c = props(6)
    phi_rad = phi * 3.141592653589793d0 / 180.0d0
    srt32 = sqrt(3.0d0 / 2.d0)
    q_phi = props(25)
    k_phi = props(26)
    if (dp_type == 1) then
       q_phi = 2.0d0 * sin(phi_rad) / (srt32 * (3.0d0 - sin(phi_rad))) * 3.d0
       k_phi = 6.0d0 * c * cos(phi_rad) / (srt32 * (3.0d0 - sin(phi_rad)))
    else if (dp_type == 2) then
       q_phi = 2.0d0 * sin(phi_rad) / (srt32 * (3.0d0 + sin(phi_rad))) * 3.d0
       k_phi = 6.0d0 * c * cos(phi_rad) / (srt32 * (3.0d0 + sin(phi_rad)))
    else if (dp_type == 3) then
       q_phi = sin(phi_rad) / sqrt(9.0d0 + 3.0d0 * sin(phi_rad) * sin(phi_rad)) / sqrt(0.5d0) * 3.d0
       k_phi = 3.0d0 * c * cos(phi_rad) / sqrt(9.0d0 + 3.0d0 * sin(phi_rad) * sin(phi_rad)) / sqrt(0.5d0)
    else if (dp_type == 4) then
       q_phi = props(5) * 3.d0
       k_phi = props(6)
    end if
    props(25) = q_phi
    props(26) = k_phi
    props(27) = 0.0d0
  end subroutine drucker_prager_prep

  ! ------------------------------------------------------------------ the whole deck
  subroutine deck_load(control_path, precision, m)
    character(*), intent(in) :: control_path
    integer(c_int32_t), intent(in) :: precision
    type(deck_model), intent(out) :: m
    character(len=LL), allocatable :: lines(:)
    character(len=LL) :: kfile, dir, card
    character(len=100) :: names(1000), sname
    integer :: nl, i, j, k, ios, n, numel, ns, sidx, iauto, iql, rk(3), iperi, isup, ifin, ilag, itype, mtype, sno
    integer :: slash, node
    real(c_double) :: dltval, qcoef, qlen, pdstime, grav(3), vel(3), nwin, props(30), bforce(3)
    real(c_double) :: e, nu, lam, mu, dmax(3), qmin
    integer :: fixity(3), body
    logical :: have_time, have_ql, have_qlen, have_rk, have_form, have_int, have_win, have_mat, have_outv, out_basic
    real(c_double), allocatable :: coo(:), win0(:), vol(:), norm_win(:), prop(:), vinit(:)
    integer(c_int32_t), allocatable :: elcon(:), part(:), set_id(:), set_ptr(:), set_node(:), hex(:)
    integer(c_int32_t), allocatable :: ebc(:), mat(:), body_id(:)

    call read_lines(control_path, lines, nl)
    slash = index(control_path, '/', back=.true.)
    dir = ''
    if (slash > 0) dir = control_path(1:slash)
    m%dir = dir

    ! control cards (PRE_CONTROL.F90:132-588) and their defaults (:589-843)
    kfile = ''
    have_time = .false.; have_ql = .false.; have_qlen = .false.; have_rk = .false.
    have_form = .false.; have_int = .false.; have_outv = .false.; out_basic = .false.
    iql = 1; qcoef = 0.001d0; qlen = 0.0d0; rk = [1, 3, 0]; iperi = 0; isup = 0; pdstime = 0.0d0
    ifin = 1; ilag = 0; itype = 2; grav = 0.0d0; m%time_output = -1.0d0; iauto = 0; dltval = 0.0d0
    do i = 1, nl
       card = lines(i)
       select case (trim(card))
       case ('*MODEL')
          kfile = adjustl(val(lines, nl, i, 1))
          kfile = kfile(1:50)
       case ('*TIME PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) m%time_end
          if (ios == 0) read(val(lines, nl, i, 2), *, iostat=ios) iauto
          if (ios == 0) read(val(lines, nl, i, 3), *, iostat=ios) dltval
          if (ios /= 0) call deck_fail('error reading the card *TIME PARAMETERS')
          have_time = .true.
       case ('*QL PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) iql, qcoef, qlen
          have_qlen = ios == 0
          if (.not. have_qlen) then
             read(val(lines, nl, i, 1), *, iostat=ios) iql, qcoef
             if (ios /= 0) call deck_fail('error reading the card *QL PARAMETERS')
          end if
          have_ql = .true.
       case ('*RK PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) rk
          if (ios /= 0) call deck_fail('error reading the card *RK PARAMETERS')
          have_rk = .true.
       case ('*PERIDYNAMICS')
          read(val(lines, nl, i, 1), *, iostat=ios) iperi
          if (ios /= 0) call deck_fail('error reading the card *PERIDYNAMICS')
       case ('*SUPPORT')
          read(val(lines, nl, i, 1), *, iostat=ios) isup
          if (ios /= 0) call deck_fail('error reading the card *SUPPORT')
       case ('*PERIODIC SEARCH')
          read(val(lines, nl, i, 1), *, iostat=ios) pdstime
          if (ios /= 0) call deck_fail('error reading the card *PERIODIC SEARCH')
       case ('*FORMULATION PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) ifin, ilag
          if (ios /= 0 .or. ifin < 0 .or. ifin > 1 .or. ilag < 0 .or. ilag > 1) &
               call deck_fail('error reading the card *FORMULATION PARAMETERS')
          have_form = .true.
       case ('*INTEGRATION PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) itype
          if (ios /= 0) call deck_fail('error reading the card *INTEGRATION PARAMETERS')
          have_int = .true.
       case ('*GRAVITY PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) grav
          if (ios /= 0) call deck_fail('error reading the card *GRAVITY PARAMETERS')
       case ('*OUTPUT PARAMETERS')
          read(val(lines, nl, i, 1), *, iostat=ios) m%time_output
          if (ios /= 0) call deck_fail('error reading the card *OUTPUT PARAMETERS')
       case ('*FEM OUTPUT')
          read(val(lines, nl, i, 1), *, iostat=ios) k
          if (ios /= 0 .or. k < 0 .or. k > 1) call deck_fail('error reading the card *FEM OUTPUT')
          m%fem_output = k == 1
       case ('*OUTPUT VARIABLES')   ! every later line names a variable (PRE_CONTROL.F90:139-214)
          have_outv = .true.
          do j = i + 1, nl
             select case (trim(lines(j)))
             case ('unformatted')
                m%unf = .true.
             case ('all')
                m%outv = .true.
             case ('basic')
                out_basic = .true.
             case default
                do k = 1, NOUTV
                   if (trim(lines(j)) == trim(out_names(k))) m%outv(k) = .true.
                end do
             end select
          end do
       end select
    end do
    if (out_basic) then           ! node_id, displacement, velocity, bodyid, eps (PRE_CONTROL.F90:208-214)
       m%outv([1, 2, 3, 7, 23]) = .true.
    end if
    if (.not. have_outv) m%outv([1, 2, 7, 8, 12, 16]) = .true.   ! defaults (PRE_CONTROL.F90:624-636)
    if (.not. have_time) call deck_fail('card *TIME PARAMETERS not found')
    ! IAUTO_TS = 1: DLT_VAL is DLT_FAC (PRE_CONTROL.F90:522-531); only the peridynamic
    ! internal force sets a step (FINT_PD.F90:690-754; FINT.F90:1250-1326 are commented out)
    if (iauto /= 0 .and. iauto /= 1) call deck_fail('choose IAUTO_TS = (0,1) in card *TIME PARAMETERS')
    if (iauto == 1 .and. iperi /= 1) &
         call deck_fail('IAUTO_TS = 1 needs *PERIDYNAMICS 1: the RKPM internal force sets no time step')
    m%auto_ts = (iauto == 1)
    m%dt = dltval
    if (m%time_output < 0.0d0) m%time_output = m%time_end * 0.01d0
    if (len_trim(kfile) == 0) call deck_fail('card *MODEL not found')

    ! model file (PRE_MODEL.F90:140-239)
    if (kfile(1:1) /= '/') kfile = trim(dir)//trim(kfile)
    call read_model_file(trim(kfile), n, coo, numel, elcon, part, ns, set_id, set_ptr, set_node)
    allocate(hex(max(numel, 1)))
    hex = 1
    i = 1
    do while (i <= numel)      ! blocks: runs of one part id; type from the block's last element
       j = i
       do while (j < numel)
          if (part(j + 1) /= part(i)) exit
          j = j + 1
       end do
       if (elcon(8 * j - 1) == elcon(8 * j)) hex(i:j) = 0
       i = j + 1
    end do
    do k = 1, size(set_node)
       if (k < set_ptr(ns + 1)) then
          if (set_node(k) < 1 .or. set_node(k) > n) call deck_fail('a node set names a node that does not exist')
       end if
    end do

    ! model cards (READ.F90:96-658)
    names = ''
    allocate(ebc(3 * n), vinit(3 * n), mat(n), norm_win(n), prop(30 * n), body_id(n))
    ebc = -1; vinit = 0.0d0; mat = -1; norm_win = 0.0d0; prop = 0.0d0; body_id = -1
    have_win = .false.; have_mat = .false.
    do i = 1, nl
       if (trim(lines(i)) == '*SET NAME') then
          read(val(lines, nl, i, 1), *, iostat=ios) sno
          if (ios == 0) read(val(lines, nl, i, 2), *, iostat=ios) sname
          if (ios /= 0 .or. sno < 1 .or. sno > 1000) call deck_fail('error reading the card *SET NAME')
          names(sno) = trim(sname)
       end if
    end do
    do i = 1, nl
       card = lines(i)
       select case (trim(card))
       case ('*BOUNDARY CONDITION')
          sidx = set_index(val(lines, nl, i, 1), ns, set_id, names)
          read(val(lines, nl, i, 2), *, iostat=ios) fixity
          if (ios /= 0) call deck_fail('error reading FIXITY in *BOUNDARY CONDITION')
          if (any(fixity == 2)) call deck_fail('prescribed nonzero displacements (fixity 2) are not on this path')
          do k = set_ptr(sidx), set_ptr(sidx + 1) - 1
             node = set_node(k)
             do j = 1, 3
                call assign_int(ebc(3 * (node - 1) + j), fixity(j), 'fixity')
             end do
          end do
       case ('*BODY DEFINITION')
          sidx = set_index(val(lines, nl, i, 1), ns, set_id, names)
          read(val(lines, nl, i, 2), *, iostat=ios) body
          if (ios == 0) read(val(lines, nl, i, 3), *, iostat=ios) bforce
          if (ios /= 0 .or. body > 1000) call deck_fail('error reading the card *BODY DEFINITION')
          do k = set_ptr(sidx), set_ptr(sidx + 1) - 1
             call assign_int(body_id(set_node(k)), body, 'body id')
          end do
       case ('*INITIAL CONDITION')
          sidx = set_index(val(lines, nl, i, 1), ns, set_id, names)
          read(val(lines, nl, i, 2), *, iostat=ios) vel
          if (ios /= 0) call deck_fail('error reading VELOCITIES in *INITIAL CONDITION')
          do k = set_ptr(sidx), set_ptr(sidx + 1) - 1
             node = set_node(k)
             do j = 1, 3
                call assign_dbl(vinit(3 * (node - 1) + j), vel(j), 'initial velocity')
             end do
          end do
       case ('*WINDOW DEFINITION')
          have_win = .true.
          sidx = set_index(val(lines, nl, i, 1), ns, set_id, names)
          read(val(lines, nl, i, 2), *, iostat=ios) nwin
          if (ios /= 0) call deck_fail('error reading NORM_WINDOW in *WINDOW DEFINITION')
          do k = set_ptr(sidx), set_ptr(sidx + 1) - 1
             call assign_dbl(norm_win(set_node(k)), nwin, 'normalised window')
          end do
       case ('*MATERIALS')
          have_mat = .true.
          sidx = set_index(val(lines, nl, i, 1), ns, set_id, names)
          read(val(lines, nl, i, 2), *, iostat=ios) mtype
          if (ios /= 0) call deck_fail('error reading MAT_TYPE in *MATERIALS')
          if (i + 11 > nl) call deck_fail('card *MATERIALS is cut short')
          do j = 0, 5            ! six rows of five after one comment line
             read(lines(i + 6 + j), *, iostat=ios) props(5 * j + 1:5 * j + 5)
             if (ios /= 0) call deck_fail('error reading a value row of *MATERIALS')
          end do
          if (mtype == 3) call drucker_prager_prep(props)
          e = props(2); nu = props(1)              ! READ.F90:489-505
          lam = e * nu / ((1.0d0 + nu) * (1.0d0 - 2.0d0 * nu))
          mu = e / (2.0d0 * (1.0d0 + nu))
          props(28) = lam; props(29) = mu; props(30) = lam + 2.0d0 * mu
          do k = set_ptr(sidx), set_ptr(sidx + 1) - 1
             node = set_node(k)
             call assign_int(mat(node), mtype, 'material type')
             prop(30 * (node - 1) + 1:30 * node) = props
          end do
       end select
    end do
    where (ebc == -1) ebc = 0                      ! unassigned dofs are free (READ.F90:532-556)
    where (body_id == -1) body_id = 1              ! READ.F90:563-569
    if (.not. have_win) then
       norm_win = 1.8d0                            ! READ.F90:603-609
    else if (any(norm_win < 0.00001d0)) then
       call deck_fail('window definitions are missing for part of the model')
    end if
    if (.not. have_mat) call deck_fail('material definitions were not found')
    if (any(mat == -1)) call deck_fail('a material definition is missing for part of the model')

    ! geometry preprocessing on the device (PRE_MODEL.F90:248-262)
    allocate(win0(3 * n), vol(n))
    call mega_check(mega_preprocess(int(n, c_int32_t), coo, int(numel, c_int32_t), elcon, hex, win0, vol, &
                                    c_null_ptr, dmax), 'mega_preprocess')

    ! QL length (PRE_CONTROL.F90:803-843): 0.1 x the smallest unscaled window
    qmin = win0(1)
    do k = 1, n
       if (win0(3 * k - 2) < qmin) qmin = win0(3 * k - 2)
       if (win0(3 * k - 1) < qmin) qmin = win0(3 * k - 1)
       if (win0(3 * k) < qmin) qmin = win0(3 * k)
    end do
    if (.not. have_ql) then
       iql = 1; qcoef = 0.001d0; qlen = qmin * 0.1d0
    else if (.not. have_qlen .and. have_rk .and. iql == 1) then
       qlen = qmin * 0.1d0
    end if

    m%n = n
    m%numel = numel
    allocate(m%x0(3 * n), m%win(3 * n), m%vol(n), m%mass(3 * n), m%prop(30 * n), m%vinit(3 * n), &
             m%ebc(3 * n), m%ebc_node(n), m%mat_type(n))
    m%x0 = coo(1:3 * n)
    do k = 1, n
       m%win(3 * k - 2:3 * k) = win0(3 * k - 2:3 * k) * norm_win(k)
       m%mass(3 * k - 2:3 * k) = vol(k) * prop(30 * (k - 1) + 3)
       m%ebc_node(k) = merge(1, 0, any(ebc(3 * k - 2:3 * k) == 1))
    end do
    m%vol = vol
    m%prop = prop
    m%ebc = ebc
    m%mat_type = mat
    m%vinit = vinit
    m%vinit_card = vinit
    allocate(m%norm_win(n), m%body_id(n), m%elcon(8 * max(numel, 1)))
    m%norm_win = norm_win
    m%body_id = body_id
    if (numel > 0) m%elcon = elcon(1:8 * numel)
    where (ebc == 1) m%vinit = 0.0d0

    m%ctrl%rk_degree = rk(1); m%ctrl%rk_cont = rk(2); m%ctrl%itype_int = itype
    m%ctrl%ql = iql; m%ctrl%ql_coef = qcoef; m%ctrl%ql_len = qlen
    m%ctrl%shsup = merge(1, 0, isup == 1)
    m%ctrl%lagrangian = ilag; m%ctrl%finite_strain = ifin
    m%ctrl%pdsearch = 1
    if (ilag == 0 .and. pdstime /= 0.0d0) m%ctrl%pdsearch = ceiling(pdstime / m%dt)   ! MAIN.F90:430
    m%ctrl%gravity = grav
    m%ctrl%dist_max = dmax
    m%ctrl%precision = precision
    m%ctrl%surrogate = 0
    m%ctrl%peridynamics = merge(1, 0, iperi == 1)
    m%ctrl%reserved = 0
  end subroutine deck_load

end module mega_deck
