! Test infrastructure only: bind(C) entry points that call the reference
! program's own subroutines (compiled from its sources into _ref/) so that the
! oracle restatement and the golden fixtures can be checked against them.
!
! Every entry point sets what the reference reads from its CONTROL module, then
! calls the reference routine with plain arrays in the reference layout
! (column-major, 1-based neighbor lists).  LOGICAL arguments arrive as int32
! 0/1 and are converted here to default LOGICAL.
!
! The bin-geometry entry point restates HANDELER.F90:187-252 (the reference
! keeps that block inline in HANDELER, which owns SAVE storage that cannot be
! reached from outside).

module ref_harness
  use iso_c_binding
  use CONTROL
  use omp_lib
  implicit none

  external :: MLS_KERNEL0, M44INV, INV3, INVERSE, DETERMINANT
  external :: HUGHES_WINGET, ROTATE_TENSOR, VON_MISES, RK1, VON_MISES_DAM, DRUCK_PRAG
  external :: GET_BINS, SOFT_SEARCH, CONSTRUCT_FINT, CONSTRUCT_FINT_PD
  external :: UDFM_SHAPE_TENSOR, DFM_SHAPE_TENSOR, ESTIMATE_MODULI

  ! bin state kept between calls, as HANDELER keeps it (HANDELER.F90:85-88)
  double precision, save :: xbin, ybin, zbin, xmin, ymin, zmin
  integer, save :: nbinsx = 0, nbinsy = 0, nbinsz = 0, nbins = 0
  integer, parameter :: max_neigh = 1000
  integer, save, allocatable :: ispace(:), jspace(:), kspace(:)
  integer, save, allocatable :: nodes_in_bin(:), nodelist_in_bin(:,:)
  ! DLT_FINT of the last CONSTRUCT_FINT_PD call (set when AUTO_TS is on)
  double precision, save :: last_dlt_fint = 0.0d0

contains

  ! AUTO_TS and DLT_FAC as PRE_CONTROL.F90:522-524 sets them (after ref_set_control)
  subroutine ref_set_auto_ts(iauto, fac) bind(C, name="ref_set_auto_ts")
    integer(c_int), value :: iauto
    real(c_double), value :: fac
    AUTO_TS = (iauto /= 0)
    DLT_FAC = fac
  end subroutine

  real(c_double) function ref_last_dlt_fint() bind(C, name="ref_last_dlt_fint")
    ref_last_dlt_fint = last_dlt_fint
  end function

  subroutine ref_set_control(rk_degree_, rk_cont_, rk_impl_, itype_int_, iql, ql_coef_, ql_len_, &
                             ishsup, ilag, ifinite, grav) bind(C, name="ref_set_control")
    integer(c_int), value :: rk_degree_, rk_cont_, rk_impl_, itype_int_, iql, ishsup, ilag, ifinite
    real(c_double), value :: ql_coef_, ql_len_
    real(c_double), intent(in) :: grav(3)
    RK_DEGREE = rk_degree_
    RK_CONT = rk_cont_
    RK_IMPL = rk_impl_
    ITYPE_INT = itype_int_
    RK_PSIZE = 4
    if (rk_degree_ == 0) RK_PSIZE = 1
    if (rk_degree_ == 2) RK_PSIZE = 10
    QL = (iql /= 0)
    QL_COEF = ql_coef_
    QL_LEN = ql_len_
    SHSUP = (ishsup /= 0)
    LLAGRANGIAN = (ilag /= 0)
    LFINITE_STRAIN = (ifinite /= 0)
    IGRAVITY = grav
    PERIDYNAMICS = .false.
    AUTO_TS = .false.
    NCORES_INPUT = 1
    PDSEARCH = 1
    call omp_set_num_threads(1)
  end subroutine

  subroutine ref_mls_kernel0(xsa, aj, ispline, phi, phix) bind(C, name="ref_mls_kernel0")
    real(c_double), value :: xsa, aj
    integer(c_int), value :: ispline
    real(c_double), intent(out) :: phi, phix
    double precision :: x, a
    integer :: isp
    logical :: iszero
    x = xsa
    a = aj
    isp = ispline
    call MLS_KERNEL0(x, a, isp, phi, phix, iszero)
  end subroutine

  subroutine ref_m44inv(a, ainv) bind(C, name="ref_m44inv")
    real(c_double), intent(in) :: a(4,4)
    real(c_double), intent(out) :: ainv(4,4)
    call M44INV(a, ainv)
  end subroutine

  subroutine ref_inv3(a, ainv) bind(C, name="ref_inv3")
    real(c_double), intent(in) :: a(3,3)
    real(c_double), intent(out) :: ainv(3,3)
    call INV3(a, ainv)
  end subroutine

  subroutine ref_inverse(a, n, ainv) bind(C, name="ref_inverse")
    integer(c_int), value :: n
    real(c_double), intent(in) :: a(n,n)
    real(c_double), intent(out) :: ainv(n,n)
    integer :: nn
    nn = n
    call INVERSE(a, nn, ainv)
  end subroutine

  subroutine ref_determinant(a, det) bind(C, name="ref_determinant")
    real(c_double), intent(in) :: a(3,3)
    real(c_double), intent(out) :: det
    double precision :: aa(3,3)
    aa = a
    call DETERMINANT(aa, det)
  end subroutine

  subroutine ref_hughes_winget(lmat, rot, strain, d) bind(C, name="ref_hughes_winget")
    real(c_double), intent(in) :: lmat(3,3)
    real(c_double), intent(out) :: rot(3,3), strain(6), d(6)
    call HUGHES_WINGET(lmat, rot, strain, d)
  end subroutine

  subroutine ref_rotate_tensor(rot, v) bind(C, name="ref_rotate_tensor")
    real(c_double), intent(in) :: rot(3,3)
    real(c_double), intent(inout) :: v(6)
    call ROTATE_TENSOR(rot, v)
  end subroutine

  subroutine ref_von_mises(stress, strain, stress_pred, state, props) bind(C, name="ref_von_mises")
    real(c_double), intent(inout) :: stress(6), state(20)
    real(c_double), intent(in) :: strain(6), stress_pred(6), props(30)
    call VON_MISES(stress, strain, stress_pred, state, props)
  end subroutine

  ! stress: in, the rotated stress of the previous step (what CONSTITUTION's
  ! LSTRESS holds at FINT.F90:797); out, the new stress
  subroutine ref_von_mises_dam(stress, total_strain, inc_strain, stress_pred, state, props) &
       bind(C, name="ref_von_mises_dam")
    real(c_double), intent(inout) :: stress(6), state(20)
    real(c_double), intent(in) :: total_strain(6), inc_strain(6), stress_pred(6), props(30)
    call VON_MISES_DAM(stress, total_strain, inc_strain, stress_pred, state, props)
  end subroutine

  subroutine ref_druck_prag(stress, total_strain, inc_strain, stress_pred, state, props) &
       bind(C, name="ref_druck_prag")
    real(c_double), intent(inout) :: stress(6), state(20)
    real(c_double), intent(in) :: total_strain(6), inc_strain(6), stress_pred(6), props(30)
    call DRUCK_PRAG(stress, total_strain, inc_strain, stress_pred, state, props)
  end subroutine

  subroutine ref_estimate_moduli(stress_inc, strain_inc, shear, bulk) bind(C, name="ref_estimate_moduli")
    real(c_double), intent(in) :: stress_inc(6), strain_inc(6)
    real(c_double), intent(inout) :: shear, bulk
    double precision :: st, bt
    logical :: nsni
    nsni = .false.
    call ESTIMATE_MODULI(stress_inc, strain_inc, st, bt, shear, bulk, nsni)
  end subroutine

  ! Shape functions at one point x for neighbor list lstack(1:ln) (RK.f90:16).
  subroutine ref_rk1(x, gnump, gcoo, gwin, lstack, ln, ebcs, self_ebc, shp, shpd) bind(C, name="ref_rk1")
    integer(c_int), value :: gnump, ln, self_ebc
    real(c_double), intent(in) :: x(3), gcoo(3,gnump), gwin(3,gnump)
    integer(c_int), intent(in) :: lstack(ln), ebcs(gnump)
    real(c_double), intent(out) :: shp(ln), shpd(3,ln)
    logical, allocatable :: lebc(:)
    logical :: lself
    integer :: n, l
    allocate(lebc(gnump))
    lebc = (ebcs /= 0)
    lself = (self_ebc /= 0)
    n = gnump
    l = ln
    call RK1(x, RK_DEGREE, RK_PSIZE, RK_CONT, RK_IMPL, gcoo, gwin, n, lstack, l, l, lebc, lself, &
             QL, QL_COEF, QL_LEN, shp, shpd, SHSUP)
    deallocate(lebc)
  end subroutine

  ! Bin geometry from positions and windows at initialisation (HANDELER.F90:187-252).
  subroutine ref_bin_geometry(gnump, gcoo, gwin, xd, yd, zd, geom, nb) bind(C, name="ref_bin_geometry")
    integer(c_int), value :: gnump
    real(c_double), intent(in) :: gcoo(3,gnump), gwin(3,gnump)
    real(c_double), value :: xd, yd, zd
    real(c_double), intent(out) :: geom(6)
    integer(c_int), intent(out) :: nb(3)
    double precision :: xmax, ymax, zmax
    integer :: i
    xbin = 0.0d0; ybin = 0.0d0; zbin = 0.0d0
    xmax = gcoo(1,1); ymax = gcoo(2,1); zmax = gcoo(3,1)
    xmin = gcoo(1,1); ymin = gcoo(2,1); zmin = gcoo(3,1)
    do i = 1, gnump
      if (gwin(1,i) > xbin) xbin = gwin(1,i)
      if (gwin(2,i) > ybin) ybin = gwin(2,i)
      if (gwin(3,i) > zbin) zbin = gwin(3,i)
      if (gwin(1,i) + xd > xbin) xbin = gwin(1,i) + xd
      if (gwin(2,i) + yd > ybin) ybin = gwin(2,i) + yd
      if (gwin(3,i) + zd > zbin) zbin = gwin(3,i) + zd
      if (gcoo(1,i) > xmax) xmax = gcoo(1,i)
      if (gcoo(2,i) > ymax) ymax = gcoo(2,i)
      if (gcoo(3,i) > zmax) zmax = gcoo(3,i)
      if (gcoo(1,i) < xmin) xmin = gcoo(1,i)
      if (gcoo(2,i) < ymin) ymin = gcoo(2,i)
      if (gcoo(3,i) < zmin) zmin = gcoo(3,i)
    end do
    xmin = xmin - xbin*10.0d0; ymin = ymin - ybin*10.0d0; zmin = zmin - zbin*10.0d0
    xmax = xmax + xbin*10.0d0; ymax = ymax + ybin*10.0d0; zmax = zmax + zbin*10.0d0
    nbinsx = ceiling((xmax - xmin)/xbin)
    nbinsy = ceiling((ymax - ymin)/ybin)
    nbinsz = ceiling((zmax - zmin)/zbin)
    nbins = nbinsx*nbinsy*nbinsz
    if (allocated(ispace)) deallocate(ispace, jspace, kspace, nodes_in_bin, nodelist_in_bin)
    allocate(ispace(gnump), jspace(gnump), kspace(gnump))
    allocate(nodes_in_bin(nbins), nodelist_in_bin(nbins, max_neigh))
    geom = (/ xmin, ymin, zmin, xbin, ybin, zbin /)
    nb = (/ nbinsx, nbinsy, nbinsz /)
  end subroutine

  ! Re-bin nodes at positions x with the stored geometry (BIN_SEARCH.F90:2).
  subroutine ref_get_bins(gnump, x, ijk) bind(C, name="ref_get_bins")
    integer(c_int), value :: gnump
    real(c_double), intent(in) :: x(3,gnump)
    integer(c_int), intent(out) :: ijk(3,gnump)
    integer :: n, i
    n = gnump
    call GET_BINS(n, x, nodes_in_bin, max_neigh, nodelist_in_bin, nbins, nbinsx, nbinsy, nbinsz, &
                  ispace, jspace, kspace, xmin, ymin, zmin, xbin, ybin, zbin)
    do i = 1, gnump
      ijk(1,i) = ispace(i); ijk(2,i) = jspace(i); ijk(3,i) = kspace(i)
    end do
  end subroutine

  ! Neighbor search at positions x against the current bins (BIN_SEARCH.F90:153).
  ! gstack must hold dim entries; returns the number of entries written in nnz.
  subroutine ref_soft_search(gnump, x, gwin, gsm_len, xd, yd, zd, gn, gstart, dim, gstack, gmaxn, nnz) &
      bind(C, name="ref_soft_search")
    integer(c_int), value :: gnump, dim
    real(c_double), intent(in) :: x(3,gnump), gwin(3,gnump), gsm_len(6,gnump)
    real(c_double), value :: xd, yd, zd
    integer(c_int), intent(out) :: gn(gnump), gstart(gnump), gstack(dim), gmaxn, nnz
    integer :: n, d, mx
    double precision :: a, b, c
    n = gnump
    d = dim
    a = xd; b = yd; c = zd
    call SOFT_SEARCH(n, x, gn, gstart, d, gstack, mx, gwin, nodes_in_bin, max_neigh, nodelist_in_bin, &
                     nbins, nbinsx, nbinsy, nbinsz, ispace, jspace, kspace, a, b, c, gsm_len)
    gmaxn = mx
    nnz = sum(gn)
  end subroutine

  ! One call of CONSTRUCT_FINT (FINT.F90:16) for material/state arrays in the
  ! reference layout.  Arrays the direct-nodal-integration path never reads are
  ! allocated here and zeroed.  linit must be 0 after the first call in a process
  ! (the reference allocates its SAVE'd GWIN0 on every LINIT call).
  subroutine ref_construct_fint(gnump, gwin, gvol, gcoo, gcoo_cur, gprop, gstate, gstress, gstrain, &
      gdinc, gdinc_tot, gmat_type, gebc, gn, gstart, dim, gstack, gshp, gdshp, gmaxn, linit, &
      fint, fext, gmax_wvel, bodyforce, body_id, gstrain_eq, gint_work, dlt) bind(C, name="ref_construct_fint")
    integer(c_int), value :: gnump, dim, gmaxn, linit
    real(c_double), intent(in) :: gwin(3,gnump), gvol(gnump), gcoo(3,gnump), gcoo_cur(3,gnump)
    real(c_double), intent(in) :: gprop(30,gnump), gdinc(3*gnump), gdinc_tot(3*gnump), bodyforce(3,gnump)
    real(c_double), intent(inout) :: gstate(20,gnump), gstress(6,gnump), gstrain(6,gnump)
    integer(c_int), intent(in) :: gmat_type(gnump), gebc(gnump), gn(gnump), gstart(gnump), gstack(dim)
    integer(c_int), intent(in) :: body_id(gnump)
    real(c_double), intent(inout) :: gshp(dim), gdshp(3,dim)
    real(c_double), intent(out) :: fint(3*gnump), fext(3*gnump), gmax_wvel(gnump), gstrain_eq(gnump)
    real(c_double), intent(inout) :: gint_work
    real(c_double), value :: dlt
    double precision, allocatable :: sm_len(:,:), sm_area(:,:), sm_vol(:), nsni_fac(:,:), h_s(:,:), s_s(:,:)
    double precision, allocatable :: ddshp(:,:), ginvk(:,:,:), dxs(:,:), dys(:,:), dzs(:,:)
    double precision, allocatable :: xm(:), ym(:), zm(:), chard(:)
    integer, allocatable :: ghost(:)
    logical, allocatable :: lebc(:)
    logical :: lin
    double precision :: dlt_fint, dl
    integer :: n, d, mx
    n = gnump; d = dim; mx = gmaxn
    allocate(sm_len(6,n), sm_area(3,n), sm_vol(n), nsni_fac(3,n), h_s(6,n), s_s(6,n))
    allocate(ddshp(6,d), ginvk(3,3,n), dxs(6,n), dys(6,n), dzs(6,n), xm(n), ym(n), zm(n), chard(n))
    allocate(ghost(n), lebc(n))
    sm_len = 0; sm_area = 0; sm_vol = 0; nsni_fac = 0; h_s = 0; s_s = 0
    ddshp = 0; ginvk = 0; dxs = 0; dys = 0; dzs = 0; xm = 0; ym = 0; zm = 0; chard = 0
    ghost = 0
    lebc = (gebc /= 0)
    lin = (linit /= 0)
    dlt_fint = 0.0d0
    dl = dlt
    call CONSTRUCT_FINT(gwin, gvol, n, gcoo, gcoo_cur, sm_len, sm_area, sm_vol, nsni_fac, &
        ghost, gprop, gstate, gstress, gstrain, h_s, s_s, gdinc, gdinc_tot, gmat_type, lebc, &
        gn, gstart, d, gstack, gshp, gdshp, ddshp, mx, ginvk, lin, &
        fint, fext, dlt_fint, chard, gmax_wvel, dxs, dys, dzs, xm, ym, zm, bodyforce, gint_work, &
        body_id, gstrain_eq, dl)
    deallocate(sm_len, sm_area, sm_vol, nsni_fac, h_s, s_s, ddshp, ginvk, dxs, dys, dzs, xm, ym, zm, chard)
    deallocate(ghost, lebc)
  end subroutine

  ! UDFM_SHAPE_TENSOR at one point (RK.f90:579-726): influence weights and the
  ! inverse of the undeformed peridynamic shape tensor.
  subroutine ref_udfm(x, gnump, gcoo, gvol, gwin, lstack, ln, ebcs, self_ebc, shp, invk) bind(C, name="ref_udfm")
    integer(c_int), value :: gnump, ln, self_ebc
    real(c_double), intent(in) :: x(3), gcoo(3,gnump), gvol(gnump), gwin(3,gnump)
    integer(c_int), intent(in) :: lstack(ln), ebcs(gnump)
    real(c_double), intent(out) :: shp(ln), invk(3,3)
    logical, allocatable :: lebc(:)
    logical :: lself
    integer :: n, l
    allocate(lebc(gnump))
    lebc = (ebcs /= 0)
    lself = (self_ebc /= 0)
    n = gnump
    l = ln
    call UDFM_SHAPE_TENSOR(x, RK_DEGREE, RK_PSIZE, RK_CONT, RK_IMPL, gcoo, gvol, gwin, n, lstack, l, l, lebc, lself, &
                           QL, QL_COEF, QL_LEN, shp, invk)
    deallocate(lebc)
  end subroutine

  ! One CONSTRUCT_FINT_PD call (FINT_PD.F90:14-759) with harness-owned
  ! neighbor lists, influence weights (gshp) and inverse shape tensors (ginvk).
  subroutine ref_construct_fint_pd(gnump, gwin, gvol, gcoo, gcoo_cur, gprop, gstate, gstress, gstrain, &
      gdinc, gdinc_tot, gmat_type, gebc, gn, gstart, dim, gstack, gshp, ginvk, gmaxn, linit, &
      fint, gmax_wvel, dlt) bind(C, name="ref_construct_fint_pd")
    integer(c_int), value :: gnump, dim, gmaxn, linit
    real(c_double), intent(in) :: gwin(3,gnump), gvol(gnump), gcoo(3,gnump), gcoo_cur(3,gnump)
    real(c_double), intent(in) :: gprop(30,gnump), gdinc(3*gnump), gdinc_tot(3*gnump)
    real(c_double), intent(inout) :: gstate(20,gnump), gstress(6,gnump), gstrain(6,gnump)
    integer(c_int), intent(in) :: gmat_type(gnump), gebc(gnump), gn(gnump), gstart(gnump), gstack(dim)
    real(c_double), intent(inout) :: gshp(dim), ginvk(3,3,gnump)
    real(c_double), intent(out) :: fint(3*gnump), gmax_wvel(gnump)
    real(c_double), value :: dlt
    double precision, allocatable :: sm_len(:,:), sm_area(:,:), sm_vol(:), nsni_fac(:,:), h_s(:,:), s_s(:,:)
    double precision, allocatable :: dshp(:,:), ddshp(:,:), dxs(:,:), dys(:,:), dzs(:,:)
    double precision, allocatable :: xm(:), ym(:), zm(:), chard(:), bod(:,:)
    integer, allocatable :: ghost(:)
    logical, allocatable :: lebc(:)
    logical :: lin
    double precision :: dlt_fint, dl, work
    integer :: n, d, mx
    n = gnump; d = dim; mx = gmaxn
    allocate(sm_len(6,n), sm_area(3,n), sm_vol(n), nsni_fac(3,n), h_s(6,n), s_s(6,n))
    allocate(dshp(3,d), ddshp(6,d), dxs(6,n), dys(6,n), dzs(6,n), xm(n), ym(n), zm(n), chard(n), bod(3,n))
    allocate(ghost(n), lebc(n))
    sm_len = 0; sm_area = 0; sm_vol = 0; nsni_fac = 0; h_s = 0; s_s = 0
    dshp = 0; ddshp = 0; dxs = 0; dys = 0; dzs = 0; xm = 0; ym = 0; zm = 0; chard = 0; bod = 0
    ghost = 0
    lebc = (gebc /= 0)
    lin = (linit /= 0)
    dlt_fint = 0.0d0
    dl = dlt
    work = 0.0d0
    call CONSTRUCT_FINT_PD(gwin, gvol, n, gcoo, gcoo_cur, sm_len, sm_area, sm_vol, nsni_fac, &
        ghost, gprop, gstate, gstress, gstrain, h_s, s_s, gdinc, gdinc_tot, gmat_type, lebc, &
        gn, gstart, d, gstack, gshp, dshp, ddshp, mx, ginvk, lin, &
        fint, dlt_fint, chard, gmax_wvel, dxs, dys, dzs, xm, ym, zm, bod, work, dl)
    last_dlt_fint = dlt_fint
    deallocate(sm_len, sm_area, sm_vol, nsni_fac, h_s, s_s, dshp, ddshp, dxs, dys, dzs, xm, ym, zm, chard, bod)
    deallocate(ghost, lebc)
  end subroutine

  ! Geometry preprocessing for an all-hex mesh (PRE_MODEL.F90:248-260):
  ! windows (GET_ADIAL), nodal volumes (GET_VOLUMES) and smoothing lengths
  ! (GET_SMOOTHING_LENGTHS) from unscaled windows, as PRE_MODEL calls them.
  subroutine ref_preprocess(numpts, numels, idelm, xyz, adial, vol, sm_len, dmax) bind(C, name="ref_preprocess")
    integer(c_int), value :: numpts, numels
    integer(c_int), intent(in) :: idelm(8,numels)
    real(c_double), intent(in) :: xyz(3,numpts)
    real(c_double), intent(out) :: adial(3,numpts), vol(numpts), sm_len(6,numpts), dmax(3)
    external :: GET_ADIAL, GET_VOLUMES, GET_SMOOTHING_LENGTHS
    integer :: typel(1000)
    integer, allocatable :: elbid(:), el(:,:)
    double precision, allocatable :: sm_area(:,:), sm_vol(:), xm(:), ym(:), zm(:)
    integer :: np, ne
    np = numpts; ne = numels
    typel = 0
    typel(1) = 2
    allocate(elbid(ne), el(8,ne), sm_area(3,np), sm_vol(np), xm(np), ym(np), zm(np))
    elbid = 1
    el = idelm
    call GET_ADIAL(adial, np, ne, el, xyz, typel, elbid)
    call GET_VOLUMES(vol, np, ne, el, xyz, typel, elbid)
    call GET_SMOOTHING_LENGTHS(vol, np, xyz, adial, sm_len, sm_area, sm_vol, dmax(1), dmax(2), dmax(3), xm, ym, zm)
    deallocate(elbid, el, sm_area, sm_vol, xm, ym, zm)
  end subroutine

end module ref_harness
