!> Fortran host driver of the tbx_h0 binding: reads one structure with its
!> flattened GFN2 tables, density and potential from a stream file, calls
!> tbx_h0_get_hamiltonian, tbx_h0_get_hamiltonian_gradient and
!> tbx_h0_get_multipole_integrals on HIP device 0 and writes the dense
!> integrals, the accumulated gradient terms and the full double-loop S, D, Q.
!> Usage: tbx_h0_driver <input> <output>
program tbx_h0_driver
   use, intrinsic :: iso_c_binding, only : c_int, c_double, c_ptr, c_associated
   use tbx_h0
   implicit none
   integer(c_int) :: hdr(8)
   integer :: nat, nid, msh, nao, nsh, ntrans, nspin, nnbr, stat1, stat2, stat3, u
   type(tbx_h0_tables) :: tab
   integer(c_int), allocatable :: id(:), ish_at(:), iao_sh(:), inl(:), nnl(:), nlat(:), nltr(:)
   real(c_double), allocatable :: xyz(:, :), trans(:, :), se(:), dsedcn(:), vao(:, :), &
      & vdp(:, :, :), vqp(:, :, :), pmat(:, :, :), xmat(:, :, :)
   real(c_double), allocatable :: s(:, :), d(:, :, :), q(:, :, :), h(:, :), dEdcn(:), &
      & grad(:, :), sigma(:, :), s2(:, :), d2(:, :, :), q2(:, :, :)
   real(c_double) :: cutoff
   character(len=512) :: fin, fout
   type(c_ptr) :: ctx

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
   read(u) hdr
   nat = hdr(1); nid = hdr(2); msh = hdr(3); nao = hdr(4)
   nsh = hdr(5); ntrans = hdr(6); nspin = hdr(7); nnbr = hdr(8)
   tab%nid = nid
   tab%msh = msh
   allocate(id(nat), xyz(3, nat), trans(3, ntrans), ish_at(nat), iao_sh(nsh))
   allocate(tab%nsh_id(nid), tab%ang(msh, nid), tab%nprim(msh, nid))
   allocate(tab%alpha(tbx_maxg, msh, nid), tab%coeff(tbx_maxg, msh, nid))
   allocate(tab%hscale(msh, msh, nid, nid), tab%shpoly(msh, nid), tab%rad(nid))
   allocate(se(nsh), dsedcn(nsh), vao(nao, nspin), vdp(3, nat, nspin), vqp(6, nat, nspin))
   allocate(pmat(nao, nao, nspin), xmat(nao, nao, nspin))
   allocate(inl(nat), nnl(nat), nlat(nnbr), nltr(nnbr))
   read(u) id, xyz, trans, cutoff, ish_at, iao_sh
   read(u) tab%nsh_id, tab%ang, tab%nprim, tab%alpha, tab%coeff, tab%hscale, tab%shpoly, tab%rad
   read(u) tab%intcut, se, dsedcn, vao, vdp, vqp, pmat, xmat
   read(u) inl, nnl, nlat, nltr
   close(u)

   ctx = tbx_h0_context(0)
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
   call tbx_h0_release(ctx)

   open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
   write(u) int([stat1, stat2, stat3], c_int)
   write(u) s, d, q, h, dEdcn, grad, sigma, s2, d2, q2
   close(u)
   if (stat1 /= 0 .or. stat2 /= 0 .or. stat3 /= 0) error stop 1
end program tbx_h0_driver
