// tbx_density.hip -- gfx950 kernels of the density matrices the nuclear
// gradient contracts, over the molecules of a batch:
//
//   get_density_matrix (scf/diag.f90:150-166)  P = C diag(focc) C^T
//   get_wdensity       (scf/diag.f90:110-143)  W = C diag(focc * eval) C^T
//   get_fermi_filling  (wavefunction/fermi.f90:31-117), one spin channel
//
// The products are the dense contractions of this path, on FP64 MFMA
// (v_mfma_f64_16x16x4_f64): one workgroup computes a 64 x 64 tile pair
// (bi <= bj) of one molecule's P from the batch's list of such tiles, each of
// its four wavefronts a 32 x 32 quarter as 2 x 2 blocks of 16 x 16, four k per
// instruction, and writes the tile and its mirror.  The operands are panels of
// C (64 rows by 16 columns for each factor) staged in LDS by the whole
// workgroup, the next panel loaded while the products run on the current one,
// so each element of C is fetched once per tile pair instead of once per
// wavefront; the tiles of a molecule run on one XCD, so C is fetched from HBM
// about once per molecule and re-read from L2.  Zero padding covers rows and
// k beyond nao.
#include <hip/hip_runtime.h>

#include "tbx_tables.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4_f64 operand layout on gfx950 (checked on the device):
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; its accumulator
// register r holds D[l / 16 + 4 r][l % 16], r = 0..3.
constexpr int KB = 16;       // k per panel
constexpr int PR = 64 + 16;  // panel row length in LDS (doubles; the pad spreads the banks)

// One 64 x 64 tile pair (bi <= bj) of one molecule's P per workgroup: the
// four wavefronts compute its 32 x 32 quarters from panels of C staged in LDS
// (64 rows x KB columns for each factor, the left one scaled by f_k), each
// panel loaded by the whole workgroup while the products run on the previous one.
__global__ void __launch_bounds__(256) k_density(TbxMols mt, const double *__restrict__ C,
                                                 const double *__restrict__ focc,
                                                 const double *__restrict__ eval,
                                                 double *__restrict__ P)
{
    __shared__ double As[KB][PR], Bs[KB][PR];
    const int w = tbx_xcd_block(blockIdx.x, gridDim.x);
    if (w >= mt.ndtile) return;  // the whole workgroup leaves together
    const int m = mt.dtiles[4 * w], bi = mt.dtiles[4 * w + 1], bj = mt.dtiles[4 * w + 2];
    const int n = mt.mol_nao[m];
    const long long d0 = mt.mol_dense[m];
    const int ao0 = mt.mol_ao[m];
    const double *Cm = C + d0;
    const int R0 = 64 * bi, C0 = 64 * bj;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int sr = 32 * (wv >> 1), sc = 32 * (wv & 1);  // the wavefront's quarter
    const int kk = tid >> 4, rr = tid & 15;             // this thread's panel column and rows
    double pa[4], pb[4];
    auto load = [&](int k0) {
        const int k = k0 + kk;
        const bool kin = k < n;
        double f = 0.0;
        if (kin) f = eval ? focc[ao0 + k] * eval[ao0 + k] : focc[ao0 + k];
        const double *col = Cm + (long long)k * n;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = R0 + rr + 16 * q, j = C0 + rr + 16 * q;
            pa[q] = (kin && i < n) ? col[i] * f : 0.0;
            pb[q] = (kin && j < n) ? col[j] : 0.0;
        }
    };
    dbl4 acc[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[s][u] = dbl4{0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int k0 = 0; k0 < n; k0 += KB) {
        __syncthreads();  // the products of the previous panel are done
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            As[kk][rr + 16 * q] = pa[q];
            Bs[kk][rr + 16 * q] = pb[q];
        }
        __syncthreads();
        if (k0 + KB < n) load(k0 + KB);
#pragma unroll
        for (int ks = 0; ks < KB; ks += 4) {
            double a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = As[ks + lk][sr + 16 * s + li];
                b[s] = Bs[ks + lk][sc + 16 * s + li];
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    acc[s][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[u], acc[s][u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = R0 + sr + 16 * s + lk + 4 * r, j = C0 + sc + 16 * u + li;
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
    if (mol->ndtile <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_density, dim3(mol->ndtile), dim3(256), 0, stream, *mol, C, focc,
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
