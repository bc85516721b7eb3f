!> Fortran host driver of the tbx_h0 binding: reads one structure with its
!> flattened GFN2 tables, density, potential and output adjoints from a stream
!> file, packs the tables with tbx_h0_pack_tables, and calls
!> tbx_h0_get_hamiltonian, tbx_h0_get_hamiltonian_gradient,
!> tbx_h0_get_multipole_integrals and tbx_h0_get_hamiltonian_adjoint on the
!> process's default context (HIP device 0).  Writes the dense integrals, the
!> accumulated gradient terms, the full double-loop S, D, Q and the adjoint's
!> dE/dxyz, dE/dse and sigma.
!> Usage: tbx_h0_driver <input> <output>
program tbx_h0_driver
   use, intrinsic :: iso_c_binding, only : c_int, c_double, c_ptr, c_associated
   use tbx_h0
   implicit none
   integer(c_int) :: hdr(8)
   integer :: nat, nid, msh, nao, nsh, ntrans, nspin, nnbr, stat0, stat1, stat2, stat3, stat4, u
   type(tbx_h0_tables) :: tab
   integer(c_int), allocatable :: id(:), ish_at(:), iao_sh(:), inl(:), nnl(:), nlat(:), nltr(:)
   integer(c_int), allocatable :: nsh_id(:), ang(:, :), nprim(:, :)
   real(c_double), allocatable :: alpha(:, :, :), coeff(:, :, :), hscale(:, :, :, :), &
      & shpoly(:, :), rad(:)
   real(c_double), allocatable :: xyz(:, :), trans(:, :), se(:), dsedcn(:), vao(:, :), &
      & vdp(:, :, :), vqp(:, :, :), pmat(:, :, :), xmat(:, :, :)
   real(c_double), allocatable :: sbar(:, :), hbar(:, :), dbar(:, :, :), qbar(:, :, :)
   real(c_double), allocatable :: s(:, :), d(:, :, :), q(:, :, :), h(:, :), dEdcn(:), &
      & grad(:, :), sigma(:, :), s2(:, :), d2(:, :, :), q2(:, :, :), agrad(:, :), adse(:), &
      & asigma(:, :)
   real(c_double) :: cutoff, intcut
   character(len=512) :: fin, fout
   type(c_ptr) :: ctx

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
   read(u) hdr
   nat = hdr(1); nid = hdr(2); msh = hdr(3); nao = hdr(4)
   nsh = hdr(5); ntrans = hdr(6); nspin = hdr(7); nnbr = hdr(8)
   allocate(id(nat), xyz(3, nat), trans(3, ntrans), ish_at(nat), iao_sh(nsh))
   allocate(nsh_id(nid), ang(msh, nid), nprim(msh, nid))
   allocate(alpha(tbx_maxg, msh, nid), coeff(tbx_maxg, msh, nid))
   allocate(hscale(msh, msh, nid, nid), shpoly(msh, nid), rad(nid))
   allocate(se(nsh), dsedcn(nsh), vao(nao, nspin), vdp(3, nat, nspin), vqp(6, nat, nspin))
   allocate(pmat(nao, nao, nspin), xmat(nao, nao, nspin))
   allocate(inl(nat), nnl(nat), nlat(nnbr), nltr(nnbr))
   allocate(sbar(nao, nao), hbar(nao, nao), dbar(3, nao, nao), qbar(6, nao, nao))
   read(u) id, xyz, trans, cutoff, ish_at, iao_sh
   read(u) nsh_id, ang, nprim, alpha, coeff, hscale, shpoly, rad
   read(u) intcut, se, dsedcn, vao, vdp, vqp, pmat, xmat
   read(u) inl, nnl, nlat, nltr
   read(u) sbar, hbar, dbar, qbar
   close(u)

   call tbx_h0_pack_tables(nsh_id, ang, nprim, alpha, coeff, intcut, hscale, shpoly, rad, tab, &
      & stat0)
   if (stat0 /= 0) then
      write(*, '(a)') "tbx_h0_driver: tbx_h0_pack_tables rejected the tables"
      error stop 3
   end if
   ctx = tbx_h0_default_context()
   if (.not. c_associated(ctx)) then
      write(*, '(a)') "tbx_h0_driver: no context: " // tbx_h0_last_error()
      error stop 2
   end if
   allocate(s(nao, nao), d(3, nao, nao), q(6, nao, nao), h(nao, nao))
   call tbx_h0_get_hamiltonian(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, nao, se, &
      & s, d, q, h, stat1, inl=inl, nnl=nnl, nlat=nlat, nltr=nltr)
   if (stat1 /= 0) write(*, '(a)') "get_hamiltonian: " // tbx_h0_last_error()
   allocate(dEdcn(nat), grad(3, nat), sigma(3, 3))
   dEdcn = 0.0_c_double
   grad = 0.0_c_double
   sigma = 0.0_c_double
   call tbx_h0_get_hamiltonian_gradient(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, nao, &
      & se, dsedcn, vao, vdp, vqp, pmat, xmat, dEdcn, grad, sigma, stat2, &
      & inl=inl, nnl=nnl, nlat=nlat, nltr=nltr)
   if (stat2 /= 0) write(*, '(a)') "get_hamiltonian_gradient: " // tbx_h0_last_error()
   allocate(s2(nao, nao), d2(3, nao, nao), q2(6, nao, nao))
   call tbx_h0_get_multipole_integrals(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, nao, &
      & s2, stat3, dpint=d2, qpint=q2)
   if (stat3 /= 0) write(*, '(a)') "get_multipole_integrals: " // tbx_h0_last_error()
   allocate(agrad(3, nat), adse(nsh), asigma(3, 3))
   agrad = 0.0_c_double
   adse = 0.0_c_double
   asigma = 0.0_c_double
   call tbx_h0_get_hamiltonian_adjoint(ctx, tab, id, xyz, trans, cutoff, ish_at, iao_sh, nao, &
      & se, sbar, hbar, dbar, qbar, agrad, adse, asigma, stat4, &
      & inl=inl, nnl=nnl, nlat=nlat, nltr=nltr)
   if (stat4 /= 0) write(*, '(a)') "get_hamiltonian_adjoint: " // tbx_h0_last_error()
   call tbx_h0_release(ctx)

   open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
   write(u) int([stat1, stat2, stat3, stat4], c_int)
   write(u) s, d, q, h, dEdcn, grad, sigma, s2, d2, q2, agrad, adse, asigma
   close(u)
   if (stat1 /= 0 .or. stat2 /= 0 .or. stat3 /= 0 .or. stat4 /= 0) error stop 1
end program tbx_h0_driver
