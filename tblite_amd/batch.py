"""Batched, device-resident GFN2 integral path over many molecules, and its
``torch.autograd.Function``.

One ``Batch`` holds the species tables and molecule offsets of ``nmol``
independent molecules (uploaded once); each call launches one kernel over all
molecules on torch's current HIP stream, so the launches order with torch work
and can be captured in a HIP graph (``torch.cuda.CUDAGraph``).

Packed layout (all float64, device):
  xyz   [3*nat]                 atoms of molecule m at atom_off[m]:atom_off[m+1]
  se    [nsh]                   self-energies (get_selfenergy) per shell
  S, H  [ndense]                molecule m: nao_m x nao_m column-major at dense_off[m]
  D     [3*ndense], Q [6*ndense]  (3|6, nao_m, nao_m) Fortran order at 3|6*dense_off[m]
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import lib as L
from . import params, structure


def _ptr(t: torch.Tensor):
    assert t.is_contiguous() and t.dtype == torch.float64
    return ct.c_void_p(t.data_ptr())


class _Description:
    """GFN2 species tables and molecule sizes of a batch, as the C-ABI's
    ``tbx_batch_desc`` (the arrays it points to are kept alive here)."""

    def __init__(self, molecules, diat_scale=None):
        numbers = [np.asarray(z, dtype=np.int32) for z, _ in molecules]
        allz = np.concatenate(numbers)
        self.nums, self.ids = params.species_of(allz)
        self.spec = params.gfn2_species(self.nums)
        if diat_scale is not None:
            # get_diat_scale of an H0 spec (xtb/spec.f90:191-211): (ksig, kpi,
            # kdel), each (nid, nid) over the batch's species (izp, jzp)
            nid = len(self.nums)
            ks = [np.asfortranarray(np.asarray(k, dtype=np.float64).reshape(nid, nid))
                  for k in diat_scale(self.nums)]
            self.spec.ksig, self.spec.kpi, self.spec.kdel = ks
        self.msh = int(self.spec.nsh_id.max())
        bas = params.Basis.new(self.ids, self.spec.nsh_id, self.spec.basis_cgto)
        self.cutoff = bas.get_cutoff()
        self.intcut = bas.intcut
        self.mol_nat = np.array([len(z) for z in numbers], dtype=np.int32)
        ang, nprim, alpha, coeff = structure.flat_cgto(self.spec, self.msh)
        self.keep = [np.ascontiguousarray(a) for a in (
            self.spec.nsh_id.astype(np.int32), ang, nprim, alpha, coeff,
            self.spec.hscale.reshape(-1, order="F"), self.spec.shpoly.reshape(-1, order="F"),
            self.spec.rad, self.mol_nat, self.ids.astype(np.int32))]
        k = self.keep
        p = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
        kd = [None] * 3
        if self.spec.ksig is not None:
            kflat = [np.ascontiguousarray(x.reshape(-1, order="F"))
                     for x in (self.spec.ksig, self.spec.kpi, self.spec.kdel)]
            self.keep += kflat
            kd = [p(x) for x in kflat]
        self.basis = L.Basis(len(self.nums), self.msh, p(k[0]), p(k[1]), p(k[2]), p(k[3]),
                             p(k[4]), self.intcut)
        self.h0 = L.H0Par(p(k[5]), p(k[6]), p(k[7]), *kd)
        self.desc = L.BatchDesc(len(molecules), p(k[8]), p(k[9]), ct.pointer(self.basis),
                                ct.pointer(self.h0), self.cutoff)


# the library's default bands (tbx_tiles' fwd_band, grad_band), for tile
# dictionaries that leave them out
DEFAULT_BANDS = {"fwd_band": 1, "grad_band": 4}


def plan_stats(molecules, tiles: dict | None = None, hint: bool = False,
               skin: float = 1.0) -> dict:
    """Work-unit plans of a batch of ``(numbers, xyz_bohr)``, built and checked
    on the host (no device): unit, pair, task and element counts, plan bytes,
    and the number of violations of what the kernels assume (``bad_pairs``,
    ``bad_elems``, ``bad_units``; all zero for a correct planner).  ``tiles``
    as in ``Batch.set_tiles`` (None: the library defaults); ``hint``: order
    the pairs near-first by the molecules' own geometry (``Batch.order_pairs``)
    and count the gradient units with a pair inside the cutoff there
    (``nunit_grad_live``)."""
    d = _Description(molecules)
    if tiles is not None:
        tiles = {**DEFAULT_BANDS, **tiles}
    t = None if tiles is None else ct.pointer(L.Tiles(**tiles))
    xyz = None
    if hint:
        xyz = np.ascontiguousarray(np.concatenate(
            [np.asarray(x, dtype=np.float64).reshape(-1) for _, x in molecules]))
    out = L.PlanStats()
    L.check(L.lib().tbx_batch_plan_stats(
        ct.byref(d.desc), t, None if xyz is None else xyz.ctypes.data_as(ct.c_void_p),
        skin, ct.byref(out)))
    return {name: int(getattr(out, name)) for name, _ in L.PlanStats._fields_}


class Batch:
    """Batched GFN2 S/D/Q/H0 + gradient for a list of ``(numbers, xyz_bohr)``.

    ``order_pairs``: order the atom pairs of the work units near-first by the
    given geometry (``order_pairs()``); it changes the speed, not the results.
    ``lattices``: optional, one entry per structure, ``None`` for a molecule
    or ``(periodic[3], lattice)`` with the lattice vectors as the rows of
    ``lattice`` (bohr); each periodic structure then sums over its lattice
    points within the larger of the integral and coordination-number cutoffs
    (``structure.get_lattice_points``, cutoff.f90:97-121).

    ``diat_scale``: optional ``f(nums) -> (ksig, kpi, kdel)``, each (nid, nid)
    over the batch's species ``nums`` (``get_diat_scale`` of an H0 spec,
    xtb/spec.f90:191-211, e.g. the CEH one, ceh/ceh.f90:1207-1242): the
    forward then builds H0 from the diatomic-frame scaled overlap
    (xtb/h0.f90:345-390); the gradient kernels refuse such a batch."""

    def __init__(self, molecules, device: int = 0, order_pairs: bool = True, skin: float = 1.0,
                 lattices=None, diat_scale=None):
        self.device = device
        d = _Description(molecules, diat_scale)
        self.nums, ids, self.spec, self.msh = d.nums, d.ids, d.spec, d.msh
        self.cutoff, self.intcut, self.mol_nat, self.ids = d.cutoff, d.intcut, d.mol_nat, d.ids
        self.xyz0 = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3)
                                    for _, x in molecules])
        self._descr = d  # keeps the tables the descriptor points to alive
        self._keep, self._basis, self._h0 = d.keep, d.basis, d.h0
        desc = d.desc
        p = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
        handle = ct.c_void_p()
        self.trans = None
        if lattices is None:
            L.check(L.lib().tbx_batch_create(L.context(device), ct.byref(desc), ct.byref(handle)))
        else:
            assert len(lattices) == len(molecules)
            rthr = max(self.cutoff, params.CN_CUTOFF)
            per = [np.zeros((1, 3)) if lat is None else
                   structure.get_lattice_points(lat[0], lat[1], rthr) for lat in lattices]
            self.mol_ntrans = np.array([len(t) for t in per], dtype=np.int32)
            self.trans = np.ascontiguousarray(np.concatenate(per))
            self.trans_off = np.concatenate([[0], np.cumsum(self.mol_ntrans)])
            L.check(L.lib().tbx_batch_create_periodic(L.context(device), ct.byref(desc),
                                                      p(self.mol_ntrans), p(self.trans),
                                                      ct.byref(handle)))
        self.handle = handle
        sz = L.BatchSizes()
        L.check(L.lib().tbx_batch_sizes_get(handle, ct.byref(sz)))
        self.nmol, self.nat, self.nsh, self.nao, self.ndense = (sz.nmol, sz.nat, sz.nsh, sz.nao,
                                                                 sz.ndense)
        n = self.nmol + 1
        self.atom_off = np.zeros(n, dtype=np.int32)
        self.sh_off = np.zeros(n, dtype=np.int32)
        self.ao_off = np.zeros(n, dtype=np.int32)
        self.dense_off = np.zeros(n, dtype=np.int64)
        L.check(L.lib().tbx_batch_offsets(handle, p(self.atom_off), p(self.sh_off),
                                          p(self.ao_off), p(self.dense_off)))
        # per-molecule views of the setup, for host-side references
        self._shell_tables()
        # the GFN2 coordination number (xtb/gfn2.f90:640): D3 covalent radii
        self.rcov = np.array([params.covalent_rad(int(z)) for z in self.nums])
        self.set_ncoord(self.rcov)
        if order_pairs:
            self.order_pairs(self.xyz0, skin)

    def _shell_tables(self):
        spec = self.spec
        self.selfenergy_sp = np.concatenate([spec.selfenergy[: spec.nsh_id[s], s]
                                             for s in self.ids])
        self.kcn_sh = np.concatenate([spec.kcn[: spec.nsh_id[s], s] for s in self.ids])
        self.sh_atom = np.repeat(np.arange(self.nat), spec.nsh_id[self.ids])
        nao_sp = np.array([sum(2 * c.ang + 1 for c in shells) for shells in spec.basis_cgto])
        self.ao_atom = np.repeat(np.arange(self.nat), nao_sp[self.ids])

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                L.lib().tbx_batch_delete(h)
            except Exception:  # interpreter shutdown
                pass

    # -- helpers -----------------------------------------------------------
    @property
    def torch_device(self):
        return torch.device("cuda", self.device)

    def tensor_xyz(self):
        return torch.tensor(self.xyz0.reshape(-1), dtype=torch.float64, device=self.torch_device)

    def selfenergy(self, cn):
        """``get_selfenergy`` over the batch (torch, differentiable in ``cn``)."""
        cn_sh = cn[torch.as_tensor(self.sh_atom, device=cn.device)]
        se0 = torch.as_tensor(self.selfenergy_sp, device=cn.device)
        kcn = torch.as_tensor(self.kcn_sh, device=cn.device)
        return se0 - kcn * cn_sh, -kcn

    def molecule_translations(self, m):
        """Translation vectors (ntr, 3) of structure ``m`` (one zero vector for
        a molecule)."""
        if self.trans is None:
            return np.zeros((1, 3))
        return self.trans[self.trans_off[m]: self.trans_off[m + 1]]

    def molecule_views(self, flat, comps=1):
        """Split a packed dense tensor into per-molecule (comps, nao, nao)
        Fortran-order views (returned as torch tensors of shape
        (nao, nao, comps) in C order = Fortran (comps, nao, nao))."""
        out = []
        for m in range(self.nmol):
            n = int(self.ao_off[m + 1] - self.ao_off[m])
            a, b = int(self.dense_off[m]), int(self.dense_off[m + 1])
            out.append(flat[comps * a: comps * b].view(n, n, comps))
        return out

    def _stream(self):
        return ct.c_void_p(torch.cuda.current_stream(self.torch_device).cuda_stream)

    # -- kernels -----------------------------------------------------------
    def forward(self, xyz, se, out=None):
        """Dense S, D, Q, H0 of every molecule (packed)."""
        dev = self.torch_device
        if out is None:
            out = (torch.empty(self.ndense, dtype=torch.float64, device=dev),
                   torch.empty(3 * self.ndense, dtype=torch.float64, device=dev),
                   torch.empty(6 * self.ndense, dtype=torch.float64, device=dev),
                   torch.empty(self.ndense, dtype=torch.float64, device=dev))
        S, D, Q, H = out
        L.check(L.lib().tbx_batch_forward(self.handle, _ptr(xyz), _ptr(se), _ptr(S), _ptr(D),
                                          _ptr(Q), _ptr(H), self._stream()))
        return S, D, Q, H

    def gradient(self, xyz, se, dsedcn, P, W, vao, vdp, vqp, out=None, S=None, D=None):
        """Plain ``get_hamiltonian_gradient`` contraction for every molecule:
        returns dEdcn [nat], gradient [3*nat] (xyz order), sigma [9*nmol].
        Given the forward's S and D at the same xyz and se, the terms linear
        in them are contracted per element (``tbx_batch_gradient_with_forward``)."""
        dev = self.torch_device
        if out is None:
            out = (torch.empty(self.nat, dtype=torch.float64, device=dev),
                   torch.empty(3 * self.nat, dtype=torch.float64, device=dev),
                   torch.empty(9 * self.nmol, dtype=torch.float64, device=dev))
        dEdcn, grad, sigma = out
        if (S is None) != (D is None):
            raise ValueError("give both S and D of the forward, or neither")
        if S is None:
            L.check(L.lib().tbx_batch_gradient(self.handle, _ptr(xyz), _ptr(se), _ptr(dsedcn),
                                               _ptr(P), _ptr(W), _ptr(vao), _ptr(vdp),
                                               _ptr(vqp), _ptr(dEdcn), _ptr(grad), _ptr(sigma),
                                               self._stream()))
        else:
            assert S.numel() == self.ndense and D.numel() == 3 * self.ndense
            L.check(L.lib().tbx_batch_gradient_with_forward(
                self.handle, _ptr(xyz), _ptr(se), _ptr(S), _ptr(D), _ptr(dsedcn), _ptr(P),
                _ptr(W), _ptr(vao), _ptr(vdp), _ptr(vqp), _ptr(dEdcn), _ptr(grad), _ptr(sigma),
                self._stream()))
        return dEdcn, grad, sigma

    def adjoint(self, xyz, se, Sb, Db, Qb, Hb, out=None, S=None, D=None):
        """Backward of ``forward``: dE/dxyz [3*nat] and dE/dse [nsh].  Given
        the forward's own S and D at the same xyz and se, the terms linear in
        them are contracted per element (``tbx_batch_adjoint_with_forward``);
        the results agree to rounding."""
        dev = self.torch_device
        if out is None:
            out = (torch.empty(3 * self.nat, dtype=torch.float64, device=dev),
                   torch.empty(self.nsh, dtype=torch.float64, device=dev))
        dxyz, dse = out
        if (S is None) != (D is None):
            raise ValueError("give both S and D of the forward, or neither")
        if S is None:
            L.check(L.lib().tbx_batch_adjoint(self.handle, _ptr(xyz), _ptr(se), _ptr(Sb),
                                              _ptr(Db), _ptr(Qb), _ptr(Hb), _ptr(dxyz),
                                              _ptr(dse), self._stream()))
        else:
            assert S.numel() == self.ndense and D.numel() == 3 * self.ndense
            L.check(L.lib().tbx_batch_adjoint_with_forward(
                self.handle, _ptr(xyz), _ptr(se), _ptr(S), _ptr(D), _ptr(Sb), _ptr(Db),
                _ptr(Qb), _ptr(Hb), _ptr(dxyz), _ptr(dse), self._stream()))
        return dxyz, dse

    # -- coordination number (mctc-lib ncoord "dexp", xtb/singlepoint.f90:224) --
    def set_ncoord(self, rcov, ka=params.CN_KA, kb=params.CN_KB, rshift=params.CN_RSHIFT,
                   cutoff=params.CN_CUTOFF):
        """Covalent radii per species (bohr) and constants of the counting
        function; the defaults are GFN2's."""
        r = np.ascontiguousarray(rcov, dtype=np.float64)
        assert r.size == len(self.nums)
        L.check(L.lib().tbx_batch_set_ncoord(self.handle, r.ctypes.data_as(ct.c_void_p),
                                             float(ka), float(kb), float(rshift), float(cutoff)))
        self._rcov_keep = r

    def ncoord(self, xyz, out=None):
        """Coordination number of every atom [nat] (device)."""
        if out is None:
            out = torch.empty(self.nat, dtype=torch.float64, device=self.torch_device)
        L.check(L.lib().tbx_batch_ncoord(self.handle, _ptr(xyz), _ptr(out), self._stream()))
        return out

    def ncoord_gradient(self, xyz, dEdcn, grad, sigma):
        """grad [3*nat] += (dcn/dxyz)^T dEdcn, sigma [9*nmol] += the strain
        part (xtb/singlepoint.f90:375-378), in place."""
        L.check(L.lib().tbx_batch_ncoord_gradient(self.handle, _ptr(xyz), _ptr(dEdcn), _ptr(grad),
                                                  _ptr(sigma), self._stream()))
        return grad, sigma

    # -- density matrices (scf/diag.f90:62-166, wavefunction/fermi.f90) -----
    def sygvd(self, H, S, eval_out=None):
        """Generalized eigenproblem H C = S C diag(eval) of every molecule:
        ``H`` (dense, packed) is overwritten with C; returns ``eval`` [nao]."""
        if eval_out is None:
            eval_out = torch.empty(self.nao, dtype=torch.float64, device=self.torch_device)
        L.check(L.lib().tbx_batch_sygvd(self.handle, _ptr(H), _ptr(S), _ptr(eval_out),
                                         self._stream()))
        return eval_out

    def fermi_filling(self, nel, kt, emo, focc=None, efermi=None):
        """``get_fermi_filling`` of one spin channel: ``nel`` [nmol] electrons,
        ``emo`` [nao] ascending; returns (focc [nao], efermi [nmol])."""
        dev = self.torch_device
        if focc is None:
            focc = torch.empty(self.nao, dtype=torch.float64, device=dev)
        if efermi is None:
            efermi = torch.empty(self.nmol, dtype=torch.float64, device=dev)
        L.check(L.lib().tbx_batch_fermi_filling(self.handle, _ptr(nel), float(kt), _ptr(emo),
                                                _ptr(focc), _ptr(efermi), self._stream()))
        return focc, efermi

    def density_matrix(self, coeff, focc, out=None):
        """``get_density_matrix``: P = C diag(focc) C^T per molecule (packed)."""
        if out is None:
            out = torch.empty(self.ndense, dtype=torch.float64, device=self.torch_device)
        L.check(L.lib().tbx_batch_get_density_matrix(self.handle, _ptr(coeff), _ptr(focc),
                                                     _ptr(out), self._stream()))
        return out

    def wdensity(self, coeff, focc, eval_, out=None):
        """Energy-weighted density of ``get_wdensity``: W = C diag(focc eval) C^T."""
        if out is None:
            out = torch.empty(self.ndense, dtype=torch.float64, device=self.torch_device)
        L.check(L.lib().tbx_batch_get_wdensity(self.handle, _ptr(coeff), _ptr(focc), _ptr(eval_),
                                               _ptr(out), self._stream()))
        return out

    # -- SCF-loop consumers of the resident integrals ------------------------
    # (scf/potential.f90:192-255, wavefunction/mulliken.f90:37-98).  Spin
    # arrays are packed spin-outermost: P and h1 [nspin*ndense], vao
    # [nspin*nao], vmp [nspin*nat*vdim]; mp is the forward's D (mpdim 3) or
    # Q (mpdim 6).

    def add_vao_to_h1(self, S, vao, h1, nspin=1):
        """``add_vao_to_h1``: h1 -= S (vao_r + vao_c)/2, in place."""
        L.check(L.lib().tbx_batch_add_vao_to_h1(self.handle, nspin, _ptr(S), _ptr(vao), _ptr(h1),
                                                self._stream()))
        return h1

    def add_vmp_to_h1(self, mp, mpdim, vmp, vdim, h1, nspin=1):
        """``add_vmp_to_h1`` with the first min(mpdim, vdim) components, in place."""
        L.check(L.lib().tbx_batch_add_vmp_to_h1(self.handle, nspin, mpdim, _ptr(mp), vdim,
                                                _ptr(vmp), _ptr(h1), self._stream()))
        return h1

    def mulliken_shell_charges(self, S, P, n0sh, nspin=1, out=None):
        """``get_mulliken_shell_charges``: qsh [nspin*nsh] (magnet representation)."""
        if out is None:
            out = torch.empty(nspin * self.nsh, dtype=torch.float64, device=self.torch_device)
        L.check(L.lib().tbx_batch_mulliken_shell_charges(self.handle, nspin, _ptr(S), _ptr(P),
                                                         _ptr(n0sh), _ptr(out), self._stream()))
        return out

    def mulliken_atomic_multipoles(self, mp, mpdim, P, nspin=1, out=None):
        """``get_mulliken_atomic_multipoles``: mpat [nspin*nat*mpdim] (magnet representation)."""
        if out is None:
            out = torch.empty(nspin * self.nat * mpdim, dtype=torch.float64,
                              device=self.torch_device)
        L.check(L.lib().tbx_batch_mulliken_atomic_multipoles(self.handle, nspin, mpdim, _ptr(mp),
                                                             _ptr(P), _ptr(out), self._stream()))
        return out

    def last_scf_timing(self):
        v = (ct.c_double * 4)()
        L.check(L.lib().tbx_batch_last_scf_timing(self.handle, v))
        return dict(zip(("add_vao_ms", "add_vmp_ms", "shell_charges_ms", "atomic_multipoles_ms"),
                        list(v)))

    def order_pairs(self, xyz=None, skin: float = 1.0):
        """Group each molecule's atom pairs closer than cutoff + ``skin`` (bohr)
        at geometry ``xyz`` ([nat, 3] or [3*nat], bohr) into the first work
        units and the others into units of their own, which the gradient
        kernels skip while their pairs stay beyond the cutoff.  Every pair is
        still screened on the device, so any geometry gives the same results;
        ``xyz=None`` restores the geometry-independent order."""
        x = None
        if xyz is not None:
            if isinstance(xyz, torch.Tensor):
                xyz = xyz.detach().cpu().numpy()
            x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1))
            assert x.size == 3 * self.nat
        L.check(L.lib().tbx_batch_order_pairs(
            self.handle, None if x is None else x.ctypes.data_as(ct.c_void_p), float(skin)))

    def tiles(self) -> dict:
        """Current tile shapes (work-unit caps and chunk lengths) of the kernels."""
        t = L.Tiles()
        L.check(L.lib().tbx_batch_get_tiles(self.handle, ct.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tiles._fields_}

    def set_tiles(self, **kw):
        """Change tile shapes (keys as returned by ``tiles()``); rebuilds the
        work units on the host and waits for pending launches of the batch."""
        cur = self.tiles()
        cur.update(kw)
        t = L.Tiles(**cur)
        L.check(L.lib().tbx_batch_set_tiles(self.handle, ct.byref(t)))

    def last_timing(self):
        v = [ct.c_double() for _ in range(4)]
        L.check(L.lib().tbx_batch_last_timing(self.handle, *[ct.byref(x) for x in v]))
        return {"forward_ms": v[0].value, "gradient_ms": v[2].value, "adjoint_ms": v[3].value}


class H0Integrals(torch.autograd.Function):
    """(xyz, se) -> (S, D, Q, H0) packed over a ``Batch``; the backward runs
    the adjoint kernel, recomputing the integrals and their derivatives
    instead of storing derivative tensors.  The outputs S and D are kept for
    the backward (``Batch.adjoint`` with S, D); modifying them in place
    before the backward makes autograd raise."""

    @staticmethod
    def forward(ctx, xyz, se, batch):
        ctx.batch = batch
        S, D, Q, H = batch.forward(xyz.detach(), se.detach())
        ctx.save_for_backward(xyz, se, S, D)
        return S, D, Q, H

    @staticmethod
    def backward(ctx, Sb, Db, Qb, Hb):
        xyz, se, S, D = ctx.saved_tensors
        b = ctx.batch
        dev = xyz.device

        def dense(t, n):
            if t is None:
                return torch.zeros(n, dtype=torch.float64, device=dev)
            return t.to(torch.float64).contiguous()

        dxyz, dse = b.adjoint(xyz.detach(), se.detach(), dense(Sb, b.ndense),
                              dense(Db, 3 * b.ndense), dense(Qb, 6 * b.ndense),
                              dense(Hb, b.ndense), S=S.detach(), D=D.detach())
        return dxyz, dse, None


def h0_integrals(batch: Batch, xyz, se):
    """Differentiable batched S, D, Q, H0 (see ``H0Integrals``)."""
    return H0Integrals.apply(xyz, se, batch)


class CoordinationNumber(torch.autograd.Function):
    """xyz -> cn over a ``Batch`` (GFN2 double-exponential counting function);
    the backward is the device chain-rule kernel (no dcn/dR array)."""

    @staticmethod
    def forward(ctx, xyz, batch):
        ctx.batch = batch
        ctx.save_for_backward(xyz)
        return batch.ncoord(xyz.detach())

    @staticmethod
    def backward(ctx, dcn):
        (xyz,) = ctx.saved_tensors
        b = ctx.batch
        dev = xyz.device
        grad = torch.zeros(3 * b.nat, dtype=torch.float64, device=dev)
        sigma = torch.zeros(9 * b.nmol, dtype=torch.float64, device=dev)
        b.ncoord_gradient(xyz.detach(), dcn.to(torch.float64).contiguous(), grad, sigma)
        return grad, None


def coordination_number(batch: Batch, xyz):
    """Differentiable coordination number (see ``CoordinationNumber``)."""
    return CoordinationNumber.apply(xyz, batch)


def gfn2_integrals(batch: Batch, xyz):
    """The differentiable GFN2 one-electron path as tblite's singlepoint runs
    it (xtb/singlepoint.f90:219-241): cn(xyz) -> get_selfenergy -> S, D, Q,
    H0.  The gradient of any function of the outputs with respect to xyz
    includes the coordination-number chain (singlepoint.f90:375-378)."""
    cn = coordination_number(batch, xyz)
    se, _ = batch.selfenergy(cn)
    return h0_integrals(batch, xyz, se.contiguous())
