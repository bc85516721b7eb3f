// tbx_scf.hip -- SCF-loop consumers of the device-resident integrals of a
// batch (gfx950, FP64):
//
//   add_vao_to_h1                  scf/potential.f90:192-219
//   add_vmp_to_h1                  scf/potential.f90:221-255
//   get_mulliken_shell_charges     wavefunction/mulliken.f90:37-62
//   get_mulliken_atomic_multipoles wavefunction/mulliken.f90:65-98
//   updown_to_magnet               wavefunction/spin.f90 (applied to the sums)
//
// Dense arrays use the batch layout: element (r, c) of molecule m at
// mol_dense[m] + c * nao_m + r, components innermost for the multipole
// integrals (D: 3, Q: 6 per element), spin outermost with stride ndense.
// Every kernel is one pass over the dense blocks: HBM-bound.
#include <hip/hip_runtime.h>

#include "tbx_tables.h"

#define SCF_THREADS 256
#define SCF_TILE 32

struct ScfArgs {
    TbxMols mt;
    int nspin;
    long long dstride;   // spin stride of dense spin arrays (P, h1): ndense
    int mpdim, vdim;     // leading dimensions of the multipole integrals / potential
    int nmp;             // components used: min(mpdim, vdim) (add_vmp_to_h1) or mpdim
    const double *S;     // [ndense]
    const double *mp;    // [mpdim * ndense]
    const double *P;     // [ndense * nspin]
    const double *vao;   // [nao_tot * nspin]
    const double *vmp;   // [vdim * nat * nspin]
    const double *n0sh;  // [nsh_tot]
    double *h1;          // [ndense * nspin], accumulated into
    double *out;         // qsh [nsh_tot * nspin] or mpat [mpdim * nat * nspin]
};

// sum_q e[q] v[q], q < NMP; 16-byte loads of e when mpdim is even (every
// element then starts on a 16-byte boundary)
template <int NMP>
__device__ __forceinline__ double mp_dot(const double *e, const double *v, int mpdim)
{
    double b = 0.0;
    if ((mpdim & 1) == 0) {
        const double2 *e2 = reinterpret_cast<const double2 *>(e);
#pragma unroll
        for (int q = 0; q + 1 < NMP; q += 2) {
            const double2 x = e2[q >> 1];
            b += x.x * v[q] + x.y * v[q + 1];
        }
        if (NMP & 1) b += e[NMP - 1] * v[NMP - 1];
    } else {
#pragma unroll
        for (int q = 0; q < NMP; ++q) b += e[q] * v[q];
    }
    return b;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// add_vao_to_h1: h1(r, c, s) -= S(r, c) (vao(r, s) + vao(c, s)) / 2, one
// wavefront per column (global AO) of every molecule.
__global__ void __launch_bounds__(SCF_THREADS) k_add_vao(ScfArgs A)
{
    const TbxMols &mt = A.mt;
    const int g = blockIdx.x * (SCF_THREADS / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nao_tot = mt.nao_tot;
    if (g >= nao_tot) return;
    const int m = mt.at_mol[mt.ao_atom[g]];
    const int a0 = mt.mol_ao[m], n = mt.mol_nao[m];
    const long long col = mt.mol_dense[m] + (long long)(g - a0) * n;
    for (int s = 0; s < A.nspin; ++s) {
        const double *v = A.vao + (long long)s * nao_tot;
        const double vc = v[g];
        double *h = A.h1 + s * A.dstride + col;
        for (int r = lane; r < n; r += 64) h[r] -= A.S[col + r] * 0.5 * (v[a0 + r] + vc);
    }
}

// add_vmp_to_h1: h1(r, c, s) -= (B(r, c) + B(c, r)) / 2 with
// B(r, c) = sum_k mp(k, r, c) vmp(k, atom(c), s): the operator of mp sits on
// the column atom.  One workgroup per pair of 32 x 32 tiles (bi, bj), bi <= bj,
// and its transpose, so that both reads are coalesced; B goes through LDS.
// NMP = min(mpdim, vdim) components.
template <int NMP>
__global__ void __launch_bounds__(SCF_THREADS) k_add_vmp(ScfArgs A)
{
    constexpr int KS = SCF_TILE / (SCF_THREADS / SCF_TILE);  // columns per thread and tile
    __shared__ double B1[SCF_TILE][SCF_TILE + 1];  // tile (bi, bj): [col][row]
    __shared__ double B2[SCF_TILE][SCF_TILE + 1];  // tile (bj, bi): [col][row]
    __shared__ double vc[2][SCF_TILE][NMP];        // potential of the column atoms of both tiles
    const TbxMols &mt = A.mt;
    const int4 tile = reinterpret_cast<const int4 *>(mt.tiles)[blockIdx.x];
    const int m = tile.x, bi = tile.y, bj = tile.z;
    const int n = mt.mol_nao[m], a0 = mt.mol_ao[m], nat = mt.nat;
    const long long dense0 = mt.mol_dense[m];
    const int rl = threadIdx.x & (SCF_TILE - 1), cl0 = threadIdx.x / SCF_TILE;
    const bool diag = bi == bj;
    const int r1 = bi * SCF_TILE + rl, r2 = bj * SCF_TILE + rl;
    for (int s = 0; s < A.nspin; ++s) {
        const double *vm = A.vmp + (long long)s * A.vdim * nat;
        double *h = A.h1 + s * A.dstride + dense0;
        if (threadIdx.x < 2 * SCF_TILE) {  // columns bj*32.. (tile 1) and bi*32.. (tile 2)
            const int t = threadIdx.x / SCF_TILE, cl = threadIdx.x % SCF_TILE;
            const int c = (t == 0 ? bj : bi) * SCF_TILE + cl;
            if (c < n) {
                const double *v = vm + (long long)A.vdim * mt.ao_atom[a0 + c];
#pragma unroll
                for (int q = 0; q < NMP; ++q) vc[t][cl][q] = v[q];
            }
        }
        __syncthreads();
        double h1a[KS], h1b[KS];  // current h1 of the two tiles, loaded with the integrals
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int cl = cl0 + k * (SCF_THREADS / SCF_TILE);
            const int c1 = bj * SCF_TILE + cl, c2 = bi * SCF_TILE + cl;
            double b1 = 0.0, b2 = 0.0;
            h1a[k] = h1b[k] = 0.0;
            if (r1 < n && c1 < n) {
                const long long e = (long long)c1 * n + r1;
                b1 = mp_dot<NMP>(A.mp + (long long)A.mpdim * (dense0 + e), vc[0][cl], A.mpdim);
                h1a[k] = h[e];
            }
            if (!diag && r2 < n && c2 < n) {
                const long long e = (long long)c2 * n + r2;
                b2 = mp_dot<NMP>(A.mp + (long long)A.mpdim * (dense0 + e), vc[1][cl], A.mpdim);
                h1b[k] = h[e];
            }
            B1[cl][rl] = b1;
            B2[cl][rl] = b2;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int cl = cl0 + k * (SCF_THREADS / SCF_TILE);
            const int c1 = bj * SCF_TILE + cl, c2 = bi * SCF_TILE + cl;
            if (r1 < n && c1 < n)
                h[(long long)c1 * n + r1] =
                    h1a[k] - 0.5 * (B1[cl][rl] + (diag ? B1[rl][cl] : B2[rl][cl]));
            if (!diag && r2 < n && c2 < n)
                h[(long long)c2 * n + r2] = h1b[k] - 0.5 * (B2[cl][rl] + B1[rl][cl]);
        }
        __syncthreads();  // B1, B2 and vc are rewritten for the next spin
    }
}

// get_mulliken_shell_charges: qsh(sh, s) = -sum_{c in sh} sum_r P(r, c, s)
// S(r, c), then updown_to_magnet and + n0sh.  One wavefront per shell.
__global__ void __launch_bounds__(SCF_THREADS) k_mulliken_shell(ScfArgs A)
{
    const TbxMols &mt = A.mt;
    const int sh = blockIdx.x * (SCF_THREADS / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nsh_tot = mt.nsh_tot;
    if (sh >= nsh_tot) return;
    const int g0 = mt.sh_aog[sh], g1 = mt.sh_aog[sh + 1];
    const int m = mt.at_mol[mt.ao_atom[g0]];
    const int a0 = mt.mol_ao[m], n = mt.mol_nao[m];
    const long long dense0 = mt.mol_dense[m];
    // the shell's columns are one contiguous run of (g1 - g0) * n elements
    const long long e0 = dense0 + (long long)(g0 - a0) * n;
    const int len = (g1 - g0) * n;
    double q[2] = {0.0, 0.0};
    for (int s = 0; s < A.nspin; ++s) {
        const double *p = A.P + s * A.dstride + e0;
        double acc = 0.0;
        for (int k = lane; k < len; k += 64) acc += p[k] * A.S[e0 + k];
        q[s] = -wave_sum(acc);
    }
    if (lane == 0) {
        if (A.nspin == 2) {
            A.out[sh] = q[0] + q[1] + A.n0sh[sh];
            A.out[sh + nsh_tot] = q[0] - q[1];
        } else {
            A.out[sh] = q[0] + A.n0sh[sh];
        }
    }
}

// get_mulliken_atomic_multipoles: mpat(k, at, s) = -sum_{c in at} sum_r
// P(r, c, s) mp(k, r, c), then updown_to_magnet.  One wavefront per atom.
__global__ void __launch_bounds__(SCF_THREADS) k_mulliken_atom(ScfArgs A)
{
    const TbxMols &mt = A.mt;
    const int at = blockIdx.x * (SCF_THREADS / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (at >= mt.nat) return;
    const int m = mt.at_mol[at];
    const int a0 = mt.mol_ao[m], n = mt.mol_nao[m];
    const int g0 = a0 + mt.at_ao[at];
    const int g1 = (at + 1 < mt.nat && mt.at_mol[at + 1] == m) ? a0 + mt.at_ao[at + 1] : a0 + n;
    const long long dense0 = mt.mol_dense[m];
    // the atom's columns are one contiguous run of (g1 - g0) * n elements
    const long long e0 = dense0 + (long long)(g0 - a0) * n;
    const int len = (g1 - g0) * n;
    double q[2][6];
    for (int s = 0; s < A.nspin; ++s) {
        const double *p = A.P + s * A.dstride + e0;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (A.mpdim == 6) {
            const double2 *e2 = reinterpret_cast<const double2 *>(A.mp + 6 * e0);
            for (int k = lane; k < len; k += 64) {
                const double pv = p[k];
                const double2 x = e2[3 * k], y = e2[3 * k + 1], z = e2[3 * k + 2];
                acc[0] += pv * x.x;
                acc[1] += pv * x.y;
                acc[2] += pv * y.x;
                acc[3] += pv * y.y;
                acc[4] += pv * z.x;
                acc[5] += pv * z.y;
            }
        } else {
            for (int k = lane; k < len; k += 64) {
                const double pv = p[k];
                const double *e = A.mp + (long long)A.mpdim * (e0 + k);
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    if (c < A.mpdim) acc[c] += pv * e[c];
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) q[s][k] = -wave_sum(acc[k]);
    }
    if (lane < A.mpdim) {
        double v0 = 0.0, v1 = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (k == lane) {
                v0 = q[0][k];
                v1 = A.nspin == 2 ? q[1][k] : 0.0;
            }
        const long long o = (long long)A.mpdim * at + lane;
        if (A.nspin == 2) {
            A.out[o] = v0 + v1;
            A.out[o + (long long)A.mpdim * mt.nat] = v0 - v1;
        } else {
            A.out[o] = v0;
        }
    }
}

extern "C" hipError_t tbx_launch_scf(int what, const TbxMols *mol, int nspin, long long dstride,
                                     int mpdim, int vdim, const double *S, const double *mp,
                                     const double *P, const double *vao, const double *vmp,
                                     const double *n0sh, double *h1, double *out,
                                     hipStream_t stream)
{
    ScfArgs a;
    a.mt = *mol;
    a.nspin = nspin;
    a.dstride = dstride;
    a.mpdim = mpdim;
    a.vdim = vdim;
    a.nmp = what == 1 ? (mpdim < vdim ? mpdim : vdim) : mpdim;
    a.S = S;
    a.mp = mp;
    a.P = P;
    a.vao = vao;
    a.vmp = vmp;
    a.n0sh = n0sh;
    a.h1 = h1;
    a.out = out;
    const int wpb = SCF_THREADS / 64;
    switch (what) {
    case 0: {  // add_vao_to_h1
        const int nao_tot = mol->nao_tot;
        if (nao_tot > 0)
            hipLaunchKernelGGL(k_add_vao, dim3((nao_tot + wpb - 1) / wpb), dim3(SCF_THREADS), 0,
                               stream, a);
        break;
    }
    case 1:  // add_vmp_to_h1
        if (mol->ntile > 0) {
            const dim3 g(mol->ntile), t(SCF_THREADS);
            switch (a.nmp) {
            case 1: hipLaunchKernelGGL(k_add_vmp<1>, g, t, 0, stream, a); break;
            case 2: hipLaunchKernelGGL(k_add_vmp<2>, g, t, 0, stream, a); break;
            case 3: hipLaunchKernelGGL(k_add_vmp<3>, g, t, 0, stream, a); break;
            case 4: hipLaunchKernelGGL(k_add_vmp<4>, g, t, 0, stream, a); break;
            case 5: hipLaunchKernelGGL(k_add_vmp<5>, g, t, 0, stream, a); break;
            default: hipLaunchKernelGGL(k_add_vmp<6>, g, t, 0, stream, a); break;
            }
        }
        break;
    case 2:  // Mulliken shell charges
        if (mol->nsh_tot > 0)
            hipLaunchKernelGGL(k_mulliken_shell, dim3((mol->nsh_tot + wpb - 1) / wpb),
                               dim3(SCF_THREADS), 0, stream, a);
        break;
    case 3:  // Mulliken atomic multipoles
        if (mol->nat > 0)
            hipLaunchKernelGGL(k_mulliken_atom, dim3((mol->nat + wpb - 1) / wpb),
                               dim3(SCF_THREADS), 0, stream, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
