// tbx_ncoord.hip -- gfx950 kernels of the GFN2 coordination number and of
// its chain rule into the nuclear gradient.
//
// cn(i) = sum_{j, T; 0 < r <= cutoff} f(r),  r = |R_i - R_j - T|,
// f(r) = 1/(1+exp(-ka (r0/r - 1))) * 1/(1+exp(-kb ((r0+rshift)/r - 1))),
// r0 = rcov(i) + rcov(j): the double-exponential counting function GFN2 uses
// (mctc-lib ncoord "dexp", created at xtb/gfn2.f90:640 and evaluated at
// xtb/singlepoint.f90:224).  The derivative kernel applies what the
// reference does with dcndr(3,nat,nat) and dcndL(3,3,nat) after
// get_hamiltonian_gradient (xtb/singlepoint.f90:375-378, two gemv calls),
// without forming either array:
//   dE/dR_i  += sum_{j,T} (dEdcn_i + dEdcn_j) f'(r) v / r,   v = R_i - R_j - T
//   sigma_m  += sum_{i in m} dEdcn_i sum_{j,T} f'(r) v v^T / r.
// One wavefront per atom i; its lanes walk the atoms j of i's molecule and
// the molecule's translations.  Every sum over j is a wavefront reduction,
// so cn and dE/dR need no atomics; sigma is summed per molecule with FP64
// atomics (six per wavefront).
#include <hip/hip_runtime.h>

#include "tbx_tables.h"

namespace {

constexpr double kEpsR2 = 2.220446049250313e-16;

struct CnPar {
    const double *rcov;  // [nsp] covalent radii (bohr)
    double ka, kb, rshift, cutoff2;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// counting function and its r-derivative
__device__ __forceinline__ void dexp_count(const CnPar &c, double r, double r0, double &f,
                                           double &df)
{
    const double a0 = r0, a1 = r0 + c.rshift;
    const double e1 = exp(-c.ka * (a0 / r - 1.0)), e2 = exp(-c.kb * (a1 / r - 1.0));
    const double s1 = 1.0 / (1.0 + e1), s2 = 1.0 / (1.0 + e2);
    const double ir2 = 1.0 / (r * r);
    // d/dr 1/(1+exp(-k (a/r - 1))) = -k a e / (r^2 (1+e)^2) = -k a e s^2 / r^2
    const double d1 = -c.ka * a0 * e1 * s1 * s1 * ir2;
    const double d2 = -c.kb * a1 * e2 * s2 * s2 * ir2;
    f = s1 * s2;
    df = d1 * s2 + s1 * d2;
}

__global__ void __launch_bounds__(256) k_ncoord(TbxMols mt, CnPar c, const double *xyz,
                                                double *cn)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= mt.nat) return;  // whole wavefronts leave together
    const int m = mt.at_mol[i];
    const int a0 = mt.mol_atom[m], a1 = mt.mol_atom[m + 1];
    const int2 tr = mt.mol_tr[m];
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    const double ri = c.rcov[mt.at_sp[i]];
    double acc = 0.0;
    for (int t = tr.x; t < tr.x + tr.y; ++t) {
        const double tx = mt.trans[3 * t], ty = mt.trans[3 * t + 1], tz = mt.trans[3 * t + 2];
        for (int j = a0 + lane; j < a1; j += 64) {
            const double vx = xi - xyz[3 * j] - tx, vy = yi - xyz[3 * j + 1] - ty,
                         vz = zi - xyz[3 * j + 2] - tz;
            const double r2 = vx * vx + vy * vy + vz * vz;
            if (r2 < kEpsR2 || r2 > c.cutoff2) continue;
            double f, df;
            dexp_count(c, sqrt(r2), ri + c.rcov[mt.at_sp[j]], f, df);
            acc += f;
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) cn[i] = acc;
}

// dE/dR (accumulated into grad) and sigma (accumulated, per molecule) from
// dE/dcn
__global__ void __launch_bounds__(256) k_ncoord_grad(TbxMols mt, CnPar c, const double *xyz,
                                                     const double *dEdcn, double *grad,
                                                     double *sigma)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= mt.nat) return;
    const int m = mt.at_mol[i];
    const int a0 = mt.mol_atom[m], a1 = mt.mol_atom[m + 1];
    const int2 tr = mt.mol_tr[m];
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    const double ri = c.rcov[mt.at_sp[i]];
    const double ei = dEdcn[i];
    double g[3] = {0.0, 0.0, 0.0};
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // xx xy yy xz yz zz
    for (int t = tr.x; t < tr.x + tr.y; ++t) {
        const double tx = mt.trans[3 * t], ty = mt.trans[3 * t + 1], tz = mt.trans[3 * t + 2];
        for (int j = a0 + lane; j < a1; j += 64) {
            const double vx = xi - xyz[3 * j] - tx, vy = yi - xyz[3 * j + 1] - ty,
                         vz = zi - xyz[3 * j + 2] - tz;
            const double r2 = vx * vx + vy * vy + vz * vz;
            if (r2 < kEpsR2 || r2 > c.cutoff2) continue;
            const double r = sqrt(r2);
            double f, df;
            dexp_count(c, r, ri + c.rcov[mt.at_sp[j]], f, df);
            const double dr = df / r;
            const double w = (ei + dEdcn[j]) * dr;
            g[0] += w * vx;
            g[1] += w * vy;
            g[2] += w * vz;
            const double ws = ei * dr;
            s[0] += ws * vx * vx;
            s[1] += ws * vx * vy;
            s[2] += ws * vy * vy;
            s[3] += ws * vx * vz;
            s[4] += ws * vy * vz;
            s[5] += ws * vz * vz;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = wave_sum(g[k]);
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) grad[3 * i + k] += g[k];
    }
    // sigma (3, 3) column-major from the packed symmetric sums
    if (lane < 9) {
        static constexpr int kPack[9] = {0, 1, 3, 1, 2, 4, 3, 4, 5};
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) v = (lane == k) ? s[kPack[k]] : v;
        if (v != 0.0) unsafeAtomicAdd(&sigma[9 * m + lane], v);
    }
}

} // namespace

extern "C" hipError_t tbx_launch_ncoord(const TbxMols *mol, const double *rcov, double ka,
                                        double kb, double rshift, double cutoff,
                                        const double *xyz, double *cn, hipStream_t stream)
{
    if (mol->nat <= 0) return hipSuccess;
    const CnPar c{rcov, ka, kb, rshift, cutoff * cutoff};
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_ncoord, dim3((mol->nat + 3) / 4), dim3(256), 0, stream, *mol, c, xyz,
                       cn);
    return hipGetLastError();
}

extern "C" hipError_t tbx_launch_ncoord_grad(const TbxMols *mol, const double *rcov, double ka,
                                             double kb, double rshift, double cutoff,
                                             const double *xyz, const double *dEdcn,
                                             double *grad, double *sigma, hipStream_t stream)
{
    if (mol->nat <= 0) return hipSuccess;
    const CnPar c{rcov, ka, kb, rshift, cutoff * cutoff};
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_ncoord_grad, dim3((mol->nat + 3) / 4), dim3(256), 0, stream, *mol, c,
                       xyz, dEdcn, grad, sigma);
    return hipGetLastError();
}
