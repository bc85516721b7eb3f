!> Fortran binding of the MI355X GFN2-xTB integral / H0 path (include/tbx.h)
!> through iso_c_binding.
!>
!> The procedures keep the argument meaning of tblite's module procedures
!>   get_selfenergy            src/tblite/xtb/h0.f90:113-217
!>   get_hamiltonian           src/tblite/xtb/h0.f90:220-429
!>   get_hamiltonian_gradient  src/tblite/xtb/h0.f90:432-666
!>   (and the adjoint of get_hamiltonian, the backward of the differentiable
!>   path, which has no reference counterpart)
!>   get_multipole_integrals, get_dipole_integrals, get_overlap (the full
!>   double-loop builders)    src/tblite/integral/native/integrals.f90:1394-1465,
!>                            994-1059, 411-470
!> with the derived types (structure_type, basis_type, tb_hamiltonian,
!> adjacency_list, potential_type) flattened to arrays.  Atom, species and
!> translation indices passed in are 1-based as in tblite and are converted to
!> the 0-based C side here; the shell/AO offsets ish_at and iao_sh are tblite's
!> 0-based offsets and pass through unchanged.  Dense outputs keep tblite's
!> column-major layout: overlap(nao,nao), dpint(3,nao,nao), qpint(6,nao,nao),
!> hamiltonian(nao,nao).  dEdcn, gradient and sigma are accumulated into.
module tbx_h0
   use, intrinsic :: iso_c_binding, only : c_int, c_double, c_ptr, c_null_ptr, c_loc, &
      & c_associated, c_char, c_f_pointer
   implicit none
   private

   public :: tbx_h0_tables, tbx_h0_context, tbx_h0_release, tbx_h0_last_error
   public :: tbx_h0_get_selfenergy, tbx_h0_get_hamiltonian, tbx_h0_get_hamiltonian_gradient
   public :: tbx_h0_get_multipole_integrals, tbx_h0_get_hamiltonian_adjoint
   public :: tbx_h0_pack_tables, tbx_h0_default_context

   integer, parameter, public :: tbx_maxg = 12, tbx_maxsh = 4

   !> Flattened basis and H0 parameters of one parametrisation
   type :: tbx_h0_tables
      integer :: nid = 0, msh = 0
      integer(c_int), allocatable :: nsh_id(:)          !< (nid)
      integer(c_int), allocatable :: ang(:, :)          !< (msh, nid)
      integer(c_int), allocatable :: nprim(:, :)        !< (msh, nid)
      real(c_double), allocatable :: alpha(:, :, :)     !< (maxg, msh, nid)
      real(c_double), allocatable :: coeff(:, :, :)     !< (maxg, msh, nid), normalised
      real(c_double), allocatable :: hscale(:, :, :, :) !< (msh, msh, nid, nid)
      real(c_double), allocatable :: shpoly(:, :)       !< (msh, nid)
      real(c_double), allocatable :: rad(:)             !< (nid), bohr
      !> diatomic-frame scaling h0%ksig, kpi, kdel (nid, nid) (xtb/h0.f90:64-65);
      !> not allocated: off (GFN2)
      real(c_double), allocatable :: ksig(:, :), kpi(:, :), kdel(:, :)
      real(c_double) :: intcut = 25.0_c_double
   end type tbx_h0_tables

   !> Context of tbx_h0_default_context (one per process, device 0)
   type(c_ptr), save :: default_ctx = c_null_ptr

   type, bind(C) :: basis_c
      integer(c_int) :: nid, msh
      type(c_ptr) :: nsh_id, ang, nprim, alpha, coeff
      real(c_double) :: intcut
   end type basis_c

   type, bind(C) :: h0par_c
      type(c_ptr) :: hscale, shpoly, rad, ksig, kpi, kdel
   end type h0par_c

   type, bind(C) :: geom_c
      integer(c_int) :: nat
      type(c_ptr) :: id, xyz
      integer(c_int) :: ntrans
      type(c_ptr) :: trans
      real(c_double) :: cutoff
   end type geom_c

   type, bind(C) :: list_c
      type(c_ptr) :: inl, nnl
      integer(c_int) :: nnbr
      type(c_ptr) :: nlat, nltr
   end type list_c

   interface
      function c_context_new(device) result(ctx) bind(C, name="tbx_context_new")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr) :: ctx
      end function c_context_new

      subroutine c_context_delete(ctx) bind(C, name="tbx_context_delete")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine c_context_delete

      function c_last_error() result(msg) bind(C, name="tbx_last_error")
         import :: c_ptr
         type(c_ptr) :: msg
      end function c_last_error

      function c_get_selfenergy(nat, id, ish_at, nsh_id, msh, se_sp, kcn, cn, se, dsedcn) &
            & result(stat) bind(C, name="tbx_get_selfenergy")
         import :: c_int, c_ptr
         integer(c_int), value :: nat, msh
         type(c_ptr), value :: id, ish_at, nsh_id, se_sp, kcn, cn, se, dsedcn
         integer(c_int) :: stat
      end function c_get_selfenergy

      function c_get_hamiltonian(ctx, geom, list, bas, h0, ish_at, iao_sh, nao, se, &
            & overlap, dpint, qpint, hamiltonian) result(stat) bind(C, name="tbx_get_hamiltonian")
         import :: c_int, c_ptr, geom_c, basis_c, h0par_c
         type(c_ptr), value :: ctx
         type(geom_c), intent(in) :: geom
         type(c_ptr), value :: list
         type(basis_c), intent(in) :: bas
         type(h0par_c), intent(in) :: h0
         type(c_ptr), value :: ish_at, iao_sh
         integer(c_int), value :: nao
         type(c_ptr), value :: se, overlap, dpint, qpint, hamiltonian
         integer(c_int) :: stat
      end function c_get_hamiltonian

      function c_get_hamiltonian_gradient(ctx, geom, list, bas, h0, ish_at, iao_sh, nao, se, &
            & dsedcn, nspin, vao, vdp, vqp, pmat, xmat, dEdcn, gradient, sigma) result(stat) &
            & bind(C, name="tbx_get_hamiltonian_gradient")
         import :: c_int, c_ptr, geom_c, basis_c, h0par_c
         type(c_ptr), value :: ctx
         type(geom_c), intent(in) :: geom
         type(c_ptr), value :: list
         type(basis_c), intent(in) :: bas
         type(h0par_c), intent(in) :: h0
         type(c_ptr), value :: ish_at, iao_sh
         integer(c_int), value :: nao
         type(c_ptr), value :: se, dsedcn
         integer(c_int), value :: nspin
         type(c_ptr), value :: vao, vdp, vqp, pmat, xmat, dEdcn, gradient, sigma
         integer(c_int) :: stat
      end function c_get_hamiltonian_gradient

      function c_get_hamiltonian_adjoint(ctx, geom, list, bas, h0, ish_at, iao_sh, nao, se, &
            & sbar, hbar, dbar, qbar, gradient, dEdse, sigma) result(stat) &
            & bind(C, name="tbx_get_hamiltonian_adjoint")
         import :: c_int, c_ptr, geom_c, basis_c, h0par_c
         type(c_ptr), value :: ctx
         type(geom_c), intent(in) :: geom
         type(c_ptr), value :: list
         type(basis_c), intent(in) :: bas
         type(h0par_c), intent(in) :: h0
         type(c_ptr), value :: ish_at, iao_sh
         integer(c_int), value :: nao
         type(c_ptr), value :: se, sbar, hbar, dbar, qbar, gradient, dEdse, sigma
         integer(c_int) :: stat
      end function c_get_hamiltonian_adjoint

      function c_get_multipole_integrals(ctx, geom, bas, ish_at, iao_sh, nao, overlap, dpint, &
            & qpint) result(stat) bind(C, name="tbx_get_multipole_integrals")
         import :: c_int, c_ptr, geom_c, basis_c
         type(c_ptr), value :: ctx
         type(geom_c), intent(in) :: geom
         type(basis_c), intent(in) :: bas
         type(c_ptr), value :: ish_at, iao_sh
         integer(c_int), value :: nao
         type(c_ptr), value :: overlap, dpint, qpint
         integer(c_int) :: stat
      end function c_get_multipole_integrals

      function c_strlen(s) result(n) bind(C, name="strlen")
         import :: c_ptr, c_int
         type(c_ptr), value :: s
         integer(c_int) :: n
      end function c_strlen
   end interface

contains

   !> Context on HIP device `device` (0-based); c_null_ptr on failure
   function tbx_h0_context(device) result(ctx)
      integer, intent(in) :: device
      type(c_ptr) :: ctx
      ctx = c_context_new(int(device, c_int))
   end function tbx_h0_context

   subroutine tbx_h0_release(ctx)
      type(c_ptr), intent(inout) :: ctx
      if (c_associated(ctx)) then
         if (c_associated(ctx, default_ctx)) default_ctx = c_null_ptr
         call c_context_delete(ctx)
      end if
      ctx = c_null_ptr
   end subroutine tbx_h0_release

   !> The process's context on HIP device 0, created on first use and kept
   !> until tbx_h0_release(ctx) is called on it (c_null_ptr on failure).  This
   !> is the context a tblite-side forwarding body passes on every call.
   function tbx_h0_default_context() result(ctx)
      type(c_ptr) :: ctx
      if (.not. c_associated(default_ctx)) default_ctx = c_context_new(0_c_int)
      ctx = default_ctx
   end function tbx_h0_default_context

   !> Fill tab from tblite's basis_type and tb_hamiltonian fields, flattened:
   !> per shell slot (ish, isp) the cgto's ang, nprim, alpha(1:nprim) and
   !> coeff(1:nprim) (bas%cgto(ish, isp), basis/type.f90:43-56), bas%nsh_id,
   !> bas%intcut, and h0%hscale(msh, msh, nid, nid), h0%shpoly(msh, nid),
   !> h0%rad(nid) (xtb/h0.f90:39-67).  Slots beyond nsh_id(isp) are ignored.
   !> With h0%do_diat_scale, pass h0%ksig, h0%kpi, h0%kdel (nid, nid) too.
   subroutine tbx_h0_pack_tables(nsh_id, ang, nprim, alpha, coeff, intcut, hscale, shpoly, &
         & rad, tab, stat, ksig, kpi, kdel)
      integer, intent(in) :: nsh_id(:), ang(:, :), nprim(:, :)
      real(c_double), intent(in) :: alpha(:, :, :), coeff(:, :, :)
      real(c_double), intent(in) :: intcut
      real(c_double), intent(in) :: hscale(:, :, :, :), shpoly(:, :), rad(:)
      type(tbx_h0_tables), intent(out) :: tab
      integer, intent(out) :: stat
      real(c_double), intent(in), optional :: ksig(:, :), kpi(:, :), kdel(:, :)
      integer :: nid, msh, isp, ish, np
      nid = size(nsh_id)
      msh = size(ang, 1)
      stat = 1
      if (msh < 1 .or. msh > tbx_maxsh .or. size(ang, 2) /= nid) return
      if (any(shape(nprim) /= [msh, nid]) .or. size(alpha, 2) /= msh .or. &
         & size(alpha, 3) /= nid .or. any(shape(coeff) /= shape(alpha))) return
      if (any(shape(hscale) /= [msh, msh, nid, nid]) .or. any(shape(shpoly) /= [msh, nid]) .or. &
         & size(rad) /= nid) return
      tab%nid = nid
      tab%msh = msh
      tab%intcut = intcut
      allocate(tab%nsh_id(nid), tab%ang(msh, nid), tab%nprim(msh, nid))
      allocate(tab%alpha(tbx_maxg, msh, nid), tab%coeff(tbx_maxg, msh, nid))
      tab%nsh_id = int(nsh_id, c_int)
      tab%ang = 0
      tab%nprim = 0
      tab%alpha = 0.0_c_double
      tab%coeff = 0.0_c_double
      do isp = 1, nid
         if (nsh_id(isp) < 1 .or. nsh_id(isp) > msh) return
         do ish = 1, nsh_id(isp)
            np = nprim(ish, isp)
            if (np < 1 .or. np > tbx_maxg .or. np > size(alpha, 1)) return
            tab%ang(ish, isp) = int(ang(ish, isp), c_int)
            tab%nprim(ish, isp) = int(np, c_int)
            tab%alpha(1:np, ish, isp) = alpha(1:np, ish, isp)
            tab%coeff(1:np, ish, isp) = coeff(1:np, ish, isp)
         end do
      end do
      tab%hscale = hscale
      tab%shpoly = shpoly
      tab%rad = rad
      if (present(ksig) .or. present(kpi) .or. present(kdel)) then
         if (.not. (present(ksig) .and. present(kpi) .and. present(kdel))) return
         if (any(shape(ksig) /= [nid, nid]) .or. any(shape(kpi) /= [nid, nid]) .or. &
            & any(shape(kdel) /= [nid, nid])) return
         tab%ksig = ksig
         tab%kpi = kpi
         tab%kdel = kdel
      end if
      stat = 0
   end subroutine tbx_h0_pack_tables

   !> The C-ABI's H0 parameter record of tab (the scaling tables when allocated)
   function h0par_of(tab) result(p)
      type(tbx_h0_tables), intent(in), target :: tab
      type(h0par_c) :: p
      p = h0par_c(c_loc(tab%hscale), c_loc(tab%shpoly), c_loc(tab%rad), c_null_ptr, &
         & c_null_ptr, c_null_ptr)
      if (allocated(tab%ksig)) then
         p%ksig = c_loc(tab%ksig)
         p%kpi = c_loc(tab%kpi)
         p%kdel = c_loc(tab%kdel)
      end if
   end function h0par_of

   !> Message of the last failing call ("" if none)
   function tbx_h0_last_error() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n, k
      p = c_last_error()
      n = 0
      if (c_associated(p)) n = c_strlen(p)
      allocate(character(len=n) :: msg)
      if (n == 0) return
      call c_f_pointer(p, s, [n])
      do k = 1, n
         msg(k:k) = s(k)
      end do
   end function tbx_h0_last_error

   !> get_selfenergy, CN branch (GFN2): se = selfenergy_sp - kcn*cn, dsedcn = -kcn
   subroutine tbx_h0_get_selfenergy(tab, id, ish_at, selfenergy_sp, kcn, cn, selfenergy, &
         & dsedcn, stat)
      type(tbx_h0_tables), intent(in), target :: tab
      integer, intent(in) :: id(:)
      integer(c_int), intent(in), target, contiguous :: ish_at(:)
      real(c_double), intent(in), target, contiguous :: selfenergy_sp(:, :), kcn(:, :), cn(:)
      real(c_double), intent(out), target, contiguous :: selfenergy(:), dsedcn(:)
      integer, intent(out) :: stat
      integer(c_int), allocatable, target :: id0(:)
      id0 = int(id - 1, c_int)
      stat = c_get_selfenergy(int(size(id), c_int), c_loc(id0), c_loc(ish_at), &
         & c_loc(tab%nsh_id), int(tab%msh, c_int), c_loc(selfenergy_sp), c_loc(kcn), c_loc(cn), &
         & c_loc(selfenergy), c_loc(dsedcn))
   end subroutine tbx_h0_get_selfenergy

   !> get_hamiltonian: dense overlap, dipole, quadrupole and H0 (intent(out)).
   !> The half list (tblite's adjacency_list: inl offsets, nnl counts, nlat
   !> 1-based neighbour atoms, nltr 1-based translations) is optional; without
   !> it the neighbours are all pairs within cutoff.
   subroutine tbx_h0_get_hamiltonian(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, nao, &
         & selfenergy, overlap, dpint, qpint, hamiltonian, stat, inl, nnl, nlat, nltr)
      type(c_ptr), intent(in) :: ctx
      type(tbx_h0_tables), intent(in), target :: tab
      integer, intent(in) :: id(:)
      real(c_double), intent(in), target, contiguous :: xyz(:, :), trans(:, :)
      real(c_double), intent(in) :: cutoff
      integer(c_int), intent(in), target, contiguous :: ish_at(:), iao_sh(:)
      integer, intent(in) :: nao
      real(c_double), intent(in), target, contiguous :: selfenergy(:)
      real(c_double), intent(out), target, contiguous :: overlap(:, :), dpint(:, :, :), &
         & qpint(:, :, :), hamiltonian(:, :)
      integer, intent(out) :: stat
      integer, intent(in), optional :: inl(:), nnl(:), nlat(:), nltr(:)
      type(geom_c) :: geom
      type(basis_c) :: bas
      type(h0par_c) :: h0
      type(list_c), target :: list
      integer(c_int), allocatable, target :: id0(:), inl0(:), nnl0(:), nlat0(:), nltr0(:)
      type(c_ptr) :: plist

      id0 = int(id - 1, c_int)
      call pack_tables(tab, bas, h0)
      geom = geom_c(int(size(id), c_int), c_loc(id0), c_loc(xyz), int(size(trans, 2), c_int), &
         & c_loc(trans), cutoff)
      plist = c_null_ptr
      if (present(inl) .and. present(nnl) .and. present(nlat) .and. present(nltr)) then
         inl0 = int(inl, c_int)
         nnl0 = int(nnl, c_int)
         nlat0 = int(nlat - 1, c_int)
         nltr0 = int(nltr - 1, c_int)
         list = list_c(c_loc(inl0), c_loc(nnl0), int(sum(nnl), c_int), c_loc(nlat0), c_loc(nltr0))
         plist = c_loc(list)
      end if
      stat = c_get_hamiltonian(ctx, geom, plist, bas, h0, c_loc(ish_at), c_loc(iao_sh), &
         & int(nao, c_int), c_loc(selfenergy), c_loc(overlap), c_loc(dpint), c_loc(qpint), &
         & c_loc(hamiltonian))
   contains
      subroutine pack_tables(t, b, p)
         type(tbx_h0_tables), intent(in), target :: t
         type(basis_c), intent(out) :: b
         type(h0par_c), intent(out) :: p
         b = basis_c(int(t%nid, c_int), int(t%msh, c_int), c_loc(t%nsh_id), c_loc(t%ang), &
            & c_loc(t%nprim), c_loc(t%alpha), c_loc(t%coeff), t%intcut)
         p = h0par_of(t)
      end subroutine pack_tables
   end subroutine tbx_h0_get_hamiltonian

   !> get_multipole_integrals / get_dipole_integrals / get_overlap: dense
   !> overlap and, when present, dpint(3,nao,nao) and qpint(6,nao,nao)
   !> (intent(out)) over every atom pair and translation within cutoff.
   subroutine tbx_h0_get_multipole_integrals(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, &
         & nao, overlap, stat, dpint, qpint)
      type(c_ptr), intent(in) :: ctx
      type(tbx_h0_tables), intent(in), target :: tab
      integer, intent(in) :: id(:)
      real(c_double), intent(in), target, contiguous :: xyz(:, :), trans(:, :)
      real(c_double), intent(in) :: cutoff
      integer(c_int), intent(in), target, contiguous :: ish_at(:), iao_sh(:)
      integer, intent(in) :: nao
      real(c_double), intent(out), target, contiguous :: overlap(:, :)
      integer, intent(out) :: stat
      real(c_double), intent(out), target, contiguous, optional :: dpint(:, :, :), qpint(:, :, :)
      type(geom_c) :: geom
      type(basis_c) :: bas
      integer(c_int), allocatable, target :: id0(:)
      type(c_ptr) :: pd, pq

      id0 = int(id - 1, c_int)
      bas = basis_c(int(tab%nid, c_int), int(tab%msh, c_int), c_loc(tab%nsh_id), c_loc(tab%ang), &
         & c_loc(tab%nprim), c_loc(tab%alpha), c_loc(tab%coeff), tab%intcut)
      geom = geom_c(int(size(id), c_int), c_loc(id0), c_loc(xyz), int(size(trans, 2), c_int), &
         & c_loc(trans), cutoff)
      pd = c_null_ptr
      pq = c_null_ptr
      if (present(dpint)) pd = c_loc(dpint)
      if (present(qpint)) pq = c_loc(qpint)
      stat = c_get_multipole_integrals(ctx, geom, bas, c_loc(ish_at), c_loc(iao_sh), &
         & int(nao, c_int), c_loc(overlap), pd, pq)
   end subroutine tbx_h0_get_multipole_integrals

   !> get_hamiltonian_gradient: contracts the integral derivatives with the
   !> density pmat and energy-weighted density xmat (nao, nao, nspin, magnet
   !> representation) and the potential (vao(nao,nspin), vdp(3,nat,nspin),
   !> vqp(6,nat,nspin)); dEdcn(nat), gradient(3,nat), sigma(3,3) accumulate.
   subroutine tbx_h0_get_hamiltonian_gradient(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, &
         & nao, selfenergy, dsedcn, vao, vdp, vqp, pmat, xmat, dEdcn, gradient, sigma, stat, &
         & inl, nnl, nlat, nltr)
      type(c_ptr), intent(in) :: ctx
      type(tbx_h0_tables), intent(in), target :: tab
      integer, intent(in) :: id(:)
      real(c_double), intent(in), target, contiguous :: xyz(:, :), trans(:, :)
      real(c_double), intent(in) :: cutoff
      integer(c_int), intent(in), target, contiguous :: ish_at(:), iao_sh(:)
      integer, intent(in) :: nao
      real(c_double), intent(in), target, contiguous :: selfenergy(:), dsedcn(:)
      real(c_double), intent(in), target, contiguous :: vao(:, :), vdp(:, :, :), vqp(:, :, :)
      real(c_double), intent(in), target, contiguous :: pmat(:, :, :), xmat(:, :, :)
      real(c_double), intent(inout), target, contiguous :: dEdcn(:), gradient(:, :), sigma(:, :)
      integer, intent(out) :: stat
      integer, intent(in), optional :: inl(:), nnl(:), nlat(:), nltr(:)
      type(geom_c) :: geom
      type(basis_c) :: bas
      type(h0par_c) :: h0
      type(list_c), target :: list
      integer(c_int), allocatable, target :: id0(:), inl0(:), nnl0(:), nlat0(:), nltr0(:)
      type(c_ptr) :: plist

      id0 = int(id - 1, c_int)
      bas = basis_c(int(tab%nid, c_int), int(tab%msh, c_int), c_loc(tab%nsh_id), c_loc(tab%ang), &
         & c_loc(tab%nprim), c_loc(tab%alpha), c_loc(tab%coeff), tab%intcut)
      h0 = h0par_of(tab)
      geom = geom_c(int(size(id), c_int), c_loc(id0), c_loc(xyz), int(size(trans, 2), c_int), &
         & c_loc(trans), cutoff)
      plist = c_null_ptr
      if (present(inl) .and. present(nnl) .and. present(nlat) .and. present(nltr)) then
         inl0 = int(inl, c_int)
         nnl0 = int(nnl, c_int)
         nlat0 = int(nlat - 1, c_int)
         nltr0 = int(nltr - 1, c_int)
         list = list_c(c_loc(inl0), c_loc(nnl0), int(sum(nnl), c_int), c_loc(nlat0), c_loc(nltr0))
         plist = c_loc(list)
      end if
      stat = c_get_hamiltonian_gradient(ctx, geom, plist, bas, h0, c_loc(ish_at), c_loc(iao_sh), &
         & int(nao, c_int), c_loc(selfenergy), c_loc(dsedcn), int(size(pmat, 3), c_int), &
         & c_loc(vao), c_loc(vdp), c_loc(vqp), c_loc(pmat), c_loc(xmat), c_loc(dEdcn), &
         & c_loc(gradient), c_loc(sigma))
   end subroutine tbx_h0_get_hamiltonian_gradient

   !> Adjoint of get_hamiltonian: with E = <sbar,S> + <hbar,H> + <dbar,D> +
   !> <qbar,Q> (sbar, hbar (nao,nao), dbar (3,nao,nao), qbar (6,nao,nao)),
   !> dE/dxyz into gradient(3,nat), dE/dselfenergy into dEdse(nsh) and the
   !> strain derivative into sigma(3,3) accumulate.
   subroutine tbx_h0_get_hamiltonian_adjoint(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, &
         & nao, selfenergy, sbar, hbar, dbar, qbar, gradient, dEdse, sigma, stat, &
         & inl, nnl, nlat, nltr)
      type(c_ptr), intent(in) :: ctx
      type(tbx_h0_tables), intent(in), target :: tab
      integer, intent(in) :: id(:)
      real(c_double), intent(in), target, contiguous :: xyz(:, :), trans(:, :)
      real(c_double), intent(in) :: cutoff
      integer(c_int), intent(in), target, contiguous :: ish_at(:), iao_sh(:)
      integer, intent(in) :: nao
      real(c_double), intent(in), target, contiguous :: selfenergy(:)
      real(c_double), intent(in), target, contiguous :: sbar(:, :), hbar(:, :), dbar(:, :, :), &
         & qbar(:, :, :)
      real(c_double), intent(inout), target, contiguous :: gradient(:, :), dEdse(:), sigma(:, :)
      integer, intent(out) :: stat
      integer, intent(in), optional :: inl(:), nnl(:), nlat(:), nltr(:)
      type(geom_c) :: geom
      type(basis_c) :: bas
      type(h0par_c) :: h0
      type(list_c), target :: list
      integer(c_int), allocatable, target :: id0(:), inl0(:), nnl0(:), nlat0(:), nltr0(:)
      type(c_ptr) :: plist

      id0 = int(id - 1, c_int)
      bas = basis_c(int(tab%nid, c_int), int(tab%msh, c_int), c_loc(tab%nsh_id), c_loc(tab%ang), &
         & c_loc(tab%nprim), c_loc(tab%alpha), c_loc(tab%coeff), tab%intcut)
      h0 = h0par_of(tab)
      geom = geom_c(int(size(id), c_int), c_loc(id0), c_loc(xyz), int(size(trans, 2), c_int), &
         & c_loc(trans), cutoff)
      plist = c_null_ptr
      if (present(inl) .and. present(nnl) .and. present(nlat) .and. present(nltr)) then
         inl0 = int(inl, c_int)
         nnl0 = int(nnl, c_int)
         nlat0 = int(nlat - 1, c_int)
         nltr0 = int(nltr - 1, c_int)
         list = list_c(c_loc(inl0), c_loc(nnl0), int(sum(nnl), c_int), c_loc(nlat0), c_loc(nltr0))
         plist = c_loc(list)
      end if
      stat = c_get_hamiltonian_adjoint(ctx, geom, plist, bas, h0, c_loc(ish_at), c_loc(iao_sh), &
         & int(nao, c_int), c_loc(selfenergy), c_loc(sbar), c_loc(hbar), c_loc(dbar), &
         & c_loc(qbar), c_loc(gradient), c_loc(dEdse), c_loc(sigma))
   end subroutine tbx_h0_get_hamiltonian_adjoint

end module tbx_h0
