"""THE ORACLE for the SCF-loop consumers of the integrals.  TEST
INFRASTRUCTURE, NOT PRODUCT CODE: only tests/ may import it, as the checker.

Plain numpy restatements, loop structure kept from the reference (tblite
0.7.0 source tree):

  add_vao_to_h1                   scf/potential.f90:192-219
  add_vmp_to_h1                   scf/potential.f90:221-255
  get_mulliken_shell_charges      wavefunction/mulliken.f90:37-62
  get_mulliken_atomic_multipoles  wavefunction/mulliken.f90:65-98
  updown_to_magnet                wavefunction/spin.f90 (updown_to_magnet_2/3)
  get_density_matrix              scf/diag.f90:150-166
  get_wdensity (one spin)         scf/diag.f90:110-143
  get_fermi_filling               wavefunction/fermi.f90:31-117
  sygvd                           lapack/sygvd.f90 (LAPACK dsygvd, here through
                                  scipy.linalg.eigh, which calls the same routine)

Arrays are Fortran-shaped numpy arrays: sint (nao, nao), mpint (nmp, nao,
nao), vao (nao, nspin), vmp (nmp, nat, nspin), h1 / pmat (nao, nao, nspin).
Indices are 0-based.
"""
import numpy as np


def add_vao_to_h1(sint, vao, h1):
    """h1 += vij for jao <= iao, mirrored (potential.f90:203-218)."""
    nao = sint.shape[0]
    for spin in range(h1.shape[2]):
        for iao in range(nao):
            jao = np.arange(iao + 1)
            vij = -sint[jao, iao] * 0.5 * (vao[jao, spin] + vao[iao, spin])
            h1[jao, iao, spin] += vij
            off = jao[:-1]
            h1[iao, off, spin] += vij[:-1]
    return h1


def add_vmp_to_h1(mpint, vmp, ao2at, h1):
    """nmp = min(size(mpint,1), size(vmp,1)) (potential.f90:236-254)."""
    nmp = min(mpint.shape[0], vmp.shape[0])
    nao = mpint.shape[1]
    for spin in range(h1.shape[2]):
        for iao in range(nao):
            iat = ao2at[iao]
            jao = np.arange(iao + 1)
            jat = ao2at[jao]
            vij = (-0.5 * np.einsum("kj,k->j", mpint[:nmp, jao, iao], vmp[:nmp, iat, spin])
                   - 0.5 * np.einsum("kj,kj->j", mpint[:nmp, iao, jao], vmp[:nmp, jat, spin]))
            h1[jao, iao, spin] += vij
            h1[iao, jao[:-1], spin] += vij[:-1]
    return h1


def updown_to_magnet(x):
    """Last axis (up, down) -> (up + down, up - down) when it has length 2."""
    if x.shape[-1] == 2:
        up, dn = x[..., 0].copy(), x[..., 1].copy()
        x[..., 0] = up + dn
        x[..., 1] = x[..., 0] - 2.0 * dn
    return x


def get_mulliken_shell_charges(smat, pmat, n0sh, iao_sh, nao_sh):
    """qsh(nsh, nspin) (mulliken.f90:44-61)."""
    nsh, nspin = len(iao_sh), pmat.shape[2]
    qsh = np.zeros((nsh, nspin))
    for spin in range(nspin):
        for ish in range(nsh):
            ii = iao_sh[ish]
            psh = 0.0
            for iao in range(nao_sh[ish]):
                psh += pmat[:, ii + iao, spin] @ smat[:, ii + iao]
            qsh[ish, spin] = -psh
    updown_to_magnet(qsh)
    qsh[:, 0] += n0sh
    return qsh


def get_mulliken_atomic_multipoles(mpmat, pmat, ao_first, nao_at):
    """mpat(nmp, nat, nspin) (mulliken.f90:72-97); ao_first / nao_at: the
    AO range of every atom (ish_at, iao_sh, nao_sh of the reference)."""
    nmp, nat, nspin = mpmat.shape[0], len(ao_first), pmat.shape[2]
    mpat = np.zeros((nmp, nat, nspin))
    for spin in range(nspin):
        for iat in range(nat):
            pat = np.zeros(nmp)
            for iao in range(ao_first[iat], ao_first[iat] + nao_at[iat]):
                pat += mpmat[:, :, iao] @ pmat[:, iao, spin]
            mpat[:, iat, spin] = -pat
    updown_to_magnet(mpat)
    return mpat


def get_density_matrix(focc, coeff):
    """scratch(jao, iao) = coeff(jao, iao) * focc(iao); pmat = scratch coeff^T
    (scf/diag.f90:150-166)."""
    scratch = coeff * focc[None, :]
    return scratch @ coeff.T


def get_wdensity(focc, eval_, coeff):
    """One spin channel of get_wdensity (scf/diag.f90:110-143): the density
    matrix of the occupations focc * eval."""
    return get_density_matrix(focc * eval_, coeff)


def sygvd(hmat, smat):
    """H C = S C diag(e), C S-orthonormal, e ascending (lapack/sygvd.f90)."""
    import scipy.linalg
    e, c = scipy.linalg.eigh(hmat, smat, driver="gvd")
    return e, c


def get_fermi_filling(nel, kt, emo):
    """Aufbau filling, then the Newton iteration of the Fermi level
    (wavefunction/fermi.f90:31-117); returns (homo, focc, e_fermi)."""
    n = len(emo)
    occ = np.zeros(n)
    homo = int(np.floor(nel))
    occ[: min(homo, n)] = 1.0
    if homo < n:
        occ[homo] = np.fmod(nel, 1.0)
    homo = homo + 1 if np.fmod(nel, 1.0) > 0.5 else homo
    e_fermi = 0.0
    if homo > 0:
        ef = 0.5 * (emo[max(homo, 1) - 1] + emo[min(homo + 1, n) - 1])
        occt = float(homo)
        thr = np.sqrt(np.finfo(float).eps)
        sqrttiny = np.sqrt(np.finfo(float).tiny)
        for _ in range(200):
            total_number = 0.0
            total_dfermi = 0.0
            for iao in range(n):
                x = (emo[iao] - ef) / kt
                fermifunct = 0.0
                if x < 50:
                    fermifunct = 1.0 / (np.exp(x) + 1.0)
                    dfermifunct = np.exp(x) / (kt * (np.exp(x) + 1.0) ** 2)
                else:
                    dfermifunct = 0.0
                occ[iao] = fermifunct
                total_number += fermifunct
                total_dfermi += dfermifunct
            change = (occt - total_number) / total_dfermi if total_dfermi > sqrttiny else 0.0
            ef += change
            if abs(occt - total_number) <= thr:
                break
        e_fermi = 0.5 * ef
    return homo, occ, e_fermi
