// tbx_density.hip -- gfx950 kernels of the density matrices the nuclear
// gradient contracts, over the molecules of a batch:
//
//   get_density_matrix (scf/diag.f90:150-166)  P = C diag(focc) C^T
//   get_wdensity       (scf/diag.f90:110-143)  W = C diag(focc * eval) C^T
//   get_fermi_filling  (wavefunction/fermi.f90:31-117), one spin channel
//
// The products are the dense contractions of this path, on FP64 MFMA
// (v_mfma_f64_16x16x4_f64): one wavefront computes a 32 x 32 tile of one
// molecule's P (the upper tile pairs bi <= bj of the batch's tile list) as
// 2 x 2 blocks of 16 x 16, four k per instruction, and writes the tile and
// its mirror.  Operands come straight from memory (C is read column-wise, 16
// consecutive rows per k; the tiles of a molecule run on one XCD, so C is
// fetched from HBM about once per molecule and re-read from L2); zero
// padding covers rows and k beyond nao.
#include <hip/hip_runtime.h>

#include "tbx_tables.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4_f64 operand layout on gfx950 (checked on the device):
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; its accumulator
// register r holds D[l / 16 + 4 r][l % 16], r = 0..3.
__global__ void __launch_bounds__(256) k_density(TbxMols mt, const double *__restrict__ C,
                                                 const double *__restrict__ focc,
                                                 const double *__restrict__ eval,
                                                 double *__restrict__ P)
{
    // four consecutive tiles per workgroup, workgroups XCD-contiguous: the tiles
    // of a molecule read its C through one L2
    const int w = tbx_xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= mt.ntile) return;  // whole wavefronts leave together
    const int m = mt.tiles[4 * w], bi = mt.tiles[4 * w + 1], bj = mt.tiles[4 * w + 2];
    const int n = mt.mol_nao[m];
    const long long d0 = mt.mol_dense[m];
    const int ao0 = mt.mol_ao[m];
    const double *Cm = C + d0;
    const int r0 = 32 * bi, c0 = 32 * bj;
    const int li = lane & 15, lk = lane >> 4;
    dbl4 acc[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[s][u] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += 4) {
        const int k = k0 + lk;
        const bool kin = k < n;
        double f = 0.0;
        if (kin) f = eval ? focc[ao0 + k] * eval[ao0 + k] : focc[ao0 + k];
        const double *col = Cm + (long long)k * n;
        double a[2], b[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int i = r0 + 16 * s + li, j = c0 + 16 * s + li;
            a[s] = (kin && i < n) ? col[i] * f : 0.0;
            b[s] = (kin && j < n) ? col[j] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int u = 0; u < 2; ++u)
                acc[s][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[u], acc[s][u], 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = r0 + 16 * s + lk + 4 * r, j = c0 + 16 * u + li;
                if (i < n && j < n) {
                    const double v = acc[s][u][r];
                    P[d0 + (long long)j * n + i] = v;
                    if (bi != bj) P[d0 + (long long)i * n + j] = v;
                }
            }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// get_fermi_filling for one spin channel of every molecule (one wavefront per
// molecule): the aufbau filling of nel electrons, then the Newton iteration
// of the Fermi level on the occupations 1/(exp((e - ef)/kt) + 1) of the
// orbital energies emo (ascending), at most 200 cycles, to sqrt(eps)
// electrons.  efermi [nmol] is half the converged level, as the reference
// returns it; a molecule with homo = 0 keeps its aufbau filling and level 0.
__global__ void __launch_bounds__(256) k_fermi(TbxMols mt, const double *__restrict__ nel,
                                               double kt, const double *__restrict__ emo,
                                               double *__restrict__ occ,
                                               double *__restrict__ efermi)
{
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (m >= mt.nmol) return;
    const int n = mt.mol_nao[m], a0 = mt.mol_ao[m];
    const double ne = nel[m];
    int homo = (int)floor(ne);
    const double frac = fmod(ne, 1.0);
    for (int k = lane; k < n; k += 64)  // get_aufbau_filling
        occ[a0 + k] = k < homo ? 1.0 : (k == homo ? frac : 0.0);
    homo = frac > 0.5 ? homo + 1 : homo;
    double ef = 0.0;
    if (homo > 0) {
        const int lo = (homo < 1 ? 1 : homo) - 1, hi = (homo + 1 < n ? homo + 1 : n) - 1;
        ef = 0.5 * (emo[a0 + lo] + emo[a0 + hi]);
        const double occt = homo;
        const double thr = 1.4901161193847656e-08;        // sqrt(epsilon(1.0_wp))
        const double sqrttiny = 1.4916681462400413e-154;  // sqrt(tiny(1.0_wp))
        for (int cycle = 0; cycle < 200; ++cycle) {
            double tn = 0.0, td = 0.0;
            for (int k = lane; k < n; k += 64) {
                const double x = (emo[a0 + k] - ef) / kt;
                double f = 0.0, df = 0.0;
                if (x < 50.0) {
                    const double e = exp(x);
                    f = 1.0 / (e + 1.0);
                    df = e / (kt * (e + 1.0) * (e + 1.0));
                }
                occ[a0 + k] = f;
                tn += f;
                td += df;
            }
            tn = wave_sum(tn);
            td = wave_sum(td);
            ef += td > sqrttiny ? (occt - tn) / td : 0.0;
            if (fabs(occt - tn) <= thr) break;
        }
        ef *= 0.5;
    }
    if (lane == 0) efermi[m] = ef;
}

} // namespace

extern "C" hipError_t tbx_launch_density(const TbxMols *mol, const double *C, const double *focc,
                                         const double *eval, double *P, hipStream_t stream)
{
    if (mol->ntile <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_density, dim3((mol->ntile + 3) / 4), dim3(256), 0, stream, *mol, C, focc,
                       eval, P);
    return hipGetLastError();
}

extern "C" hipError_t tbx_launch_fermi(const TbxMols *mol, const double *nel, double kt,
                                       const double *emo, double *occ, double *efermi,
                                       hipStream_t stream)
{
    if (mol->nmol <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fermi, dim3((mol->nmol + 3) / 4), dim3(256), 0, stream, *mol, nel, kt,
                       emo, occ, efermi);
    return hipGetLastError();
}
