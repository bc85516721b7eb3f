// tbx_kernels.hip -- gfx950 kernels of the GFN2-xTB integral / H0 path.
//
// Forward (get_hamiltonian, xtb/h0.f90:220-429): one workgroup owns the dense
// column panel of one atom i (its <= 9 spherical columns over all rows of its
// molecule).  Rows are walked in windows of consecutive atoms j.  Phase A: one
// lane per Cartesian function pair (c_j, c_i) accumulates S, D, Q (operator on
// i, the column atom) and the H0-weighted overlap over primitive pairs and
// lattice images into LDS.  Phase B: one lane per spherical element (mu, nu)
// applies the Cartesian->spherical transform (integral/trafo.f90:34-44, p in
// y,z,x order), removes the quadrupole trace and writes S, D, Q, H0 columns
// straight to HBM -- including the zero rows of atoms beyond the cutoff, so
// every dense element is written exactly once with column-contiguous stores
// and no memset pass is needed.
//
// Gradient / adjoint (get_hamiltonian_gradient, xtb/h0.f90:432-666): one
// workgroup per atom i walks the atoms j < i (the reference's half list) in
// windows.  One lane per Cartesian pair first folds the spherical weights of
// its block (from P, W and the potentials, or from the adjoints of S, H0, D,
// Q) into Cartesian weights through the transpose of the transform, then
// recomputes the primitive integrals and their R_i derivatives and contracts
// them on the fly -- no derivative tensor is ever stored.  Per-pair results go
// to atom j / shells of j through LDS and then global FP64 atomics.
//
// All arithmetic is FP64.
#include <hip/hip_runtime.h>

#include "tbx_tables.h"

#define MAXG 12
#define NTHREADS 256
#define TW_FWD 512   // Cartesian-pair tasks per forward window (LDS: 11 doubles each)
#define TW_GRAD 256  // Cartesian-pair tasks per gradient window
#define MAXSH 4

static constexpr double kS3 = 1.7320508075688772935;
static constexpr double kSqrtPi3 = 5.5683279968317078453;  // pi^(3/2)
static constexpr double kEps = 2.220446049250313e-16;

// Cartesian -> spherical: sph(l, m) = sum_k w * cart(l, c_k), CCA Cartesian
// order; l=1 (y, z, x); l=2 m=-2..2: sqrt3 xy, sqrt3 yz, zz-(xx+yy)/2,
// sqrt3 xz, sqrt3/2 (xx-yy).
__constant__ int kSphN[3][5] = {{1, 0, 0, 0, 0}, {1, 1, 1, 0, 0}, {1, 1, 3, 1, 2}};
__constant__ int kSphC[3][5][3] = {{{0}},
                                   {{1}, {2}, {0}},
                                   {{1}, {4}, {5, 0, 3}, {2}, {0, 3}}};
__constant__ double kSphW[3][5][3] = {
    {{1.0}},
    {{1.0}, {1.0}, {1.0}},
    {{kS3}, {kS3}, {1.0, -0.5, -0.5}, {kS3}, {0.5 * kS3, -0.5 * kS3}}};
// Transpose of the same map: Cartesian c -> (m, w) pairs.
__constant__ int kCartN[3][6] = {{1, 0, 0, 0, 0, 0}, {1, 1, 1, 0, 0, 0}, {2, 1, 1, 2, 1, 1}};
__constant__ int kCartM[3][6][2] = {{{0}},
                                    {{2}, {0}, {1}},
                                    {{2, 4}, {0}, {3}, {2, 4}, {1}, {2}}};
__constant__ double kCartW[3][6][2] = {
    {{1.0}},
    {{1.0}, {1.0}, {1.0}},
    {{-0.5, 0.5 * kS3}, {kS3}, {kS3}, {-0.5, -0.5 * kS3}, {kS3}, {1.0}}};

// ---------------------------------------------------------------------------
// 1D Gaussian moment factors.  In the product-centre coordinate y = x - P the
// primitive pair is exp(-eab y^2) times polynomials in y; odd moments vanish,
// even ones are (2n-1)!! h^n with h = 1/(2 eab).

struct Mom {
    double m0, m2, m4, m6, m8;
};

__device__ __forceinline__ Mom make_mom(double h)
{
    Mom M;
    M.m0 = 1.0;
    M.m2 = h;
    M.m4 = 3.0 * h * h;
    M.m6 = 15.0 * h * h * h;
    M.m8 = 105.0 * h * h * h * h;
    return M;
}

__device__ __forceinline__ double mom(const Mom &M, int t)
{
    switch (t) {
    case 0: return M.m0;
    case 2: return M.m2;
    case 4: return M.m4;
    case 6: return M.m6;
    case 8: return M.m8;
    default: return 0.0;
    }
}

// (y + p)^n for n <= 2 as coefficients c[0..2]
__device__ __forceinline__ void pow2_poly(int n, double p, double c[3])
{
    c[0] = (n == 0) ? 1.0 : ((n == 1) ? p : p * p);
    c[1] = (n == 0) ? 0.0 : ((n == 1) ? 1.0 : 2.0 * p);
    c[2] = (n == 2) ? 1.0 : 0.0;
}

// integral of sum_k P[k] y^k (y + c)^m, m = 0, 1, 2; N coefficients
template <int N>
__device__ __forceinline__ void mom3(const double *P, double c, const Mom &M, double &i0,
                                     double &i1, double &i2)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double mk = mom(M, k), mk1 = mom(M, k + 1), mk2 = mom(M, k + 2);
        a0 += P[k] * mk;
        a1 += P[k] * (mk1 + c * mk);
        a2 += P[k] * (mk2 + 2.0 * c * mk1 + c * c * mk);
    }
    i0 = a0;
    i1 = a1;
    i2 = a2;
}

template <int N>
__device__ __forceinline__ void mom12(const double *P, double c, const Mom &M, double &i1,
                                      double &i2)
{
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double mk = mom(M, k), mk1 = mom(M, k + 1), mk2 = mom(M, k + 2);
        a1 += P[k] * (mk1 + c * mk);
        a2 += P[k] * (mk2 + 2.0 * c * mk1 + c * c * mk);
    }
    i1 = a1;
    i2 = a2;
}

// Forward 1D factors of one direction: V[m] = int (y+pa)^(a+m) (y+pb)^b, m=0..2
__device__ __forceinline__ void fwd1d(int a, int b, double pa, double pb, const Mom &M,
                                      double V[3])
{
    double A[3], B[3], P[5];
    pow2_poly(a, pa, A);
    pow2_poly(b, pb, B);
#pragma unroll
    for (int k = 0; k < 5; ++k) P[k] = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) P[x + y] += A[x] * B[y];
    mom3<5>(P, pa, M, V[0], V[1], V[2]);
}

// Gradient 1D factors of one direction (derivative w.r.t. centre i, exponent ai):
//   Vi[m] = int (y+pa)^(a+m) (y+pb)^b          (operator on i)
//   Vj[m] = int (y+pa)^a (y+pb)^(b+m)          (operator on j), Vj[0] = Vi[0]
//   Gi[m] = d/dA_x of Vi[m] (function and operator move with centre i)
//   Gj[m] = d/dA_x of Vj[m]
__device__ __forceinline__ void grad1d(int a, int b, double pa, double pb, double ai,
                                       const Mom &M, double Vi[3], double Vj[3], double Gi[3],
                                       double Gj[3])
{
    double A[3], B[3], P[5], G[4], PG[6];
    pow2_poly(a, pa, A);
    pow2_poly(b, pb, B);
#pragma unroll
    for (int k = 0; k < 5; ++k) P[k] = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) P[x + y] += A[x] * B[y];
    // G = d/dA [(x-A)^a e^{-ai (x-A)^2}] / e = 2 ai (y+pa)^(a+1) - a (y+pa)^(a-1)
    double ta = 2.0 * ai, fa = (double)a;
    double Am0 = (a == 2) ? pa : ((a == 1) ? 1.0 : 0.0);
    double Am1 = (a == 2) ? 1.0 : 0.0;
    G[0] = ta * (pa * A[0]) - fa * Am0;
    G[1] = ta * (A[0] + pa * A[1]) - fa * Am1;
    G[2] = ta * (A[1] + pa * A[2]);
    G[3] = ta * A[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) PG[k] = 0.0;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) PG[x + y] += G[x] * B[y];
    mom3<5>(P, pa, M, Vi[0], Vi[1], Vi[2]);
    Vj[0] = Vi[0];
    mom12<5>(P, pb, M, Vj[1], Vj[2]);
    double g1, g2;
    mom3<6>(PG, pa, M, Gi[0], g1, g2);
    // the operator (x - A)^m moves with A as well: - m (x-A)^(m-1)
    Gi[1] = g1 - Vi[0];
    Gi[2] = g2 - 2.0 * Vi[1];
    Gj[0] = Gi[0];
    mom12<6>(PG, pb, M, Gj[1], Gj[2]);
}

__device__ __forceinline__ void unpack_exp(int e, int x[3])
{
    x[0] = e & 15;
    x[1] = (e >> 4) & 15;
    x[2] = (e >> 8) & 15;
}

// quadrupole component q -> moment exponent per direction (xx xy yy xz yz zz)
static constexpr int kQm[6][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};

// ---------------------------------------------------------------------------
// Forward: dense S, D, Q, H0 column panels.

struct FwdArgs {
    TbxSpecies sp;
    TbxMols mol;
    const double *xyz;   // [3*nat]
    const double *se;    // [nsh]
    const int *order;    // [nat] panel order (heaviest first)
    double *S, *D, *Q, *H;
    const int *nnl;      // optional [nat]: expected half-list counts (NULL: no check)
    int *flag;           // set to 1 on a list mismatch
};

__global__ void __launch_bounds__(NTHREADS) k_forward(FwdArgs A)
{
    __shared__ double cart[TW_FWD * 11];
    __shared__ int s_count;
    const TbxSpecies &sp = A.sp;
    const TbxMols &mt = A.mol;
    const int i = A.order[blockIdx.x];
    const int m = mt.at_mol[i];
    const int a0 = mt.mol_atom[m], a1 = mt.mol_atom[m + 1];
    const int nao = mt.mol_nao[m];
    const int spi = mt.at_sp[i];
    const int nci = sp.ncart[spi], naoi = sp.nao[spi];
    const double xi = A.xyz[3 * i], yi = A.xyz[3 * i + 1], zi = A.xyz[3 * i + 2];
    const long long dense0 = mt.mol_dense[m];
    const int cart_m = mt.mol_cart[m];
    const int ncart_mol = mt.mol_cart[m + 1] - cart_m;
    const int ao_m = mt.mol_ao[m];
    const int iao0 = mt.at_ao[i];
    const int rows_cap = TW_FWD / nci;  // Cartesian rows per window
    if (threadIdx.x == 0) s_count = 0;

    int c0 = 0;
    while (c0 < ncart_mol) {
        // window = atoms j whose Cartesian rows fit in [c0, c0 + rows_cap)
        const int jfirst = mt.cart_atom[cart_m + c0];
        int lo = jfirst, hi = a1;  // last atom with at_cart <= c0 + rows_cap - ncart(j) -> bsearch on start
        // largest jl in [jfirst, a1) with at_cart[jl] + ncart[jl] <= c0 + rows_cap
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            int end = mt.at_cart[mid] + sp.ncart[mt.at_sp[mid]];
            if (end <= c0 + rows_cap) lo = mid; else hi = mid;
        }
        const int jlast = lo;
        const int c1 = mt.at_cart[jlast] + sp.ncart[mt.at_sp[jlast]];
        const int ntask = (c1 - c0) * nci;
        __syncthreads();

        // Phase A: Cartesian pair integrals into LDS
        for (int t = threadIdx.x; t < ntask; t += NTHREADS) {
            const int cjl = t / nci, ci = t - cjl * nci;
            const int j = mt.cart_atom[cart_m + c0 + cjl];
            const int spj = mt.at_sp[j];
            const int cj = c0 + cjl - mt.at_cart[j];
            const int slj = sp.cart_sh[spj * TBX_MAXCART + cj];
            const int sli = sp.cart_sh[spi * TBX_MAXCART + ci];
            int ej[3], ei[3];
            unpack_exp(sp.cart_exp[spj * TBX_MAXCART + cj], ej);
            unpack_exp(sp.cart_exp[spi * TBX_MAXCART + ci], ei);
            const int slotj = slj + sp.msh * spj, sloti = sli + sp.msh * spi;
            const int npj = sp.nprim[slotj], npi = sp.nprim[sloti];
            const double dx0 = xi - A.xyz[3 * j], dy0 = yi - A.xyz[3 * j + 1],
                         dz0 = zi - A.xyz[3 * j + 2];
            const double hsc = sp.hscale[slj + sp.msh * (sli + sp.msh * (spj + sp.nsp * spi))];
            const double kpi = sp.shpoly[sloti], kpj = sp.shpoly[slotj];
            const double radij = sp.rad[spi] + sp.rad[spj];
            double acc[11];
#pragma unroll
            for (int q = 0; q < 11; ++q) acc[q] = 0.0;
            int nimg = 0;
            for (int itr = 0; itr < mt.ntrans; ++itr) {
                const double vx = dx0 - mt.trans[3 * itr], vy = dy0 - mt.trans[3 * itr + 1],
                             vz = dz0 - mt.trans[3 * itr + 2];
                const double r2 = vx * vx + vy * vy + vz * vz;
                double wH;
                if (r2 < kEps) {
                    if (j != i) continue;  // coinciding distinct atoms: not in the list
                    wH = 1.0;             // on-site block: H = S (se_i + se_j)/2
                } else {
                    if (r2 > sp.cutoff2) continue;
                    ++nimg;
                    const double rr = sqrt(sqrt(r2) / radij);
                    wH = hsc * (1.0 + kpi * rr) * (1.0 + kpj * rr);
                }
                const double vec[3] = {vx, vy, vz};
                for (int ip = 0; ip < npi; ++ip) {
                    const double ai = sp.alpha[sloti * MAXG + ip];
                    const double ci_ = sp.coeff[sloti * MAXG + ip];
                    for (int jp = 0; jp < npj; ++jp) {
                        const double aj = sp.alpha[slotj * MAXG + jp];
                        const double eab = ai + aj;
                        const double oab = 1.0 / eab;
                        const double est = ai * aj * r2 * oab;
                        if (est > sp.intcut) continue;
                        const double so = sqrt(oab);
                        const double cc = exp(-est) * kSqrtPi3 * so * so * so * ci_ *
                                          sp.coeff[slotj * MAXG + jp];
                        const Mom M = make_mom(0.5 * oab);
                        double V[3][3];
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            fwd1d(ei[k], ej[k], -vec[k] * aj * oab, vec[k] * ai * oab, M, V[k]);
                        const double s = V[0][0] * V[1][0] * V[2][0];
                        acc[0] += cc * s;
                        acc[1] += cc * V[0][1] * V[1][0] * V[2][0];
                        acc[2] += cc * V[0][0] * V[1][1] * V[2][0];
                        acc[3] += cc * V[0][0] * V[1][0] * V[2][1];
#pragma unroll
                        for (int q = 0; q < 6; ++q)
                            acc[4 + q] += cc * V[0][kQm[q][0]] * V[1][kQm[q][1]] * V[2][kQm[q][2]];
                        acc[10] += wH * cc * s;
                    }
                }
            }
            if (A.nnl && ci == 0 && cj == 0 && j <= i && nimg > 0)
                atomicAdd(&s_count, nimg);
#pragma unroll
            for (int q = 0; q < 11; ++q) cart[t * 11 + q] = acc[q];
        }
        __syncthreads();

        // Phase B: spherical transform and column writes of rows [r0, r1)
        const int r0 = mt.at_ao[jfirst];
        const int r1 = mt.at_ao[jlast] + sp.nao[mt.at_sp[jlast]];
        const int nrow = r1 - r0;
        for (int e = threadIdx.x; e < nrow * naoi; e += NTHREADS) {
            const int col = e / nrow, rl = e - col * nrow;
            const int mu = r0 + rl;
            const int j = mt.ao_atom[ao_m + mu];
            const int spj = mt.at_sp[j];
            const int aoj = mu - mt.at_ao[j];
            const int shj = sp.ao_sh[spj * TBX_MAXAO + aoj], mj = sp.ao_m[spj * TBX_MAXAO + aoj];
            const int shi = sp.ao_sh[spi * TBX_MAXAO + col], mi = sp.ao_m[spi * TBX_MAXAO + col];
            const int lj = sp.ang[shj + sp.msh * spj], li = sp.ang[shi + sp.msh * spi];
            const int cbase_j = mt.at_cart[j] - c0 + sp.sh_cart[shj + sp.msh * spj];
            const int cbase_i = sp.sh_cart[shi + sp.msh * spi];
            double v[11];
#pragma unroll
            for (int q = 0; q < 11; ++q) v[q] = 0.0;
            const int nj = kSphN[lj][mj], ni = kSphN[li][mi];
            for (int a = 0; a < nj; ++a) {
                const int rowc = cbase_j + kSphC[lj][mj][a];
                const double wa = kSphW[lj][mj][a];
                for (int b = 0; b < ni; ++b) {
                    const double w = wa * kSphW[li][mi][b];
                    const double *src = &cart[(rowc * nci + cbase_i + kSphC[li][mi][b]) * 11];
#pragma unroll
                    for (int q = 0; q < 11; ++q) v[q] += w * src[q];
                }
            }
            const double tr = 0.5 * (v[4] + v[6] + v[9]);
            const int nu = iao0 + col;
            const long long idx = dense0 + (long long)nu * nao + mu;
            const double hij = 0.5 * (A.se[mt.at_sh[i] + shi] + A.se[mt.at_sh[j] + shj]);
            A.S[idx] = v[0];
            A.H[idx] = hij * v[10];
            double *dp = A.D + 3 * idx;
            dp[0] = v[1];
            dp[1] = v[2];
            dp[2] = v[3];
            double *qp = A.Q + 6 * idx;
            qp[0] = 1.5 * v[4] - tr;
            qp[1] = 1.5 * v[5];
            qp[2] = 1.5 * v[6] - tr;
            qp[3] = 1.5 * v[7];
            qp[4] = 1.5 * v[8];
            qp[5] = 1.5 * v[9] - tr;
        }
        c0 = c1;
    }
    if (A.nnl) {
        __syncthreads();
        if (threadIdx.x == 0 && s_count != A.nnl[i]) atomicOr(A.flag, 1);
    }
    (void)a0;
}

// ---------------------------------------------------------------------------
// Gradient / adjoint contraction.

enum { GRAD_PLAIN = 0, GRAD_ADJOINT = 1 };

struct GradArgs {
    TbxSpecies sp;
    TbxMols mol;
    const double *xyz, *se;
    const int *order;
    // plain: P, W dense per molecule (spin planes at +spin*pstride), vao [nao_tot*nspin],
    // vdp [3*nat*nspin], vqp [6*nat*nspin]
    const double *P, *W, *vao, *vdp, *vqp;
    int nspin;
    long long pstride;   // elements between spin planes of P (and W)
    long long vstride;   // elements between spin planes of vao
    // adjoint: dense adjoints of S, D, Q, H0
    const double *Sb, *Db, *Qb, *Hb;
    // outputs (accumulated with atomics; caller zeroes)
    double *grad;   // [3*nat]
    double *dse;    // [nsh]  dE/d(selfenergy)
    double *sigma;  // [9*nmol]
    const int *nnl; // optional [nat] list check
    int *flag;
};

template <int MODE>
__global__ void __launch_bounds__(NTHREADS) k_gradient(GradArgs A)
{
    __shared__ double accJ[TW_GRAD][3 + MAXSH];  // per window atom: dG (3), dse of its shells
    __shared__ double accI[3 + MAXSH + 9];       // atom i: dG, dse, sigma
    __shared__ int s_count;
    const TbxSpecies &sp = A.sp;
    const TbxMols &mt = A.mol;
    const int i = A.order[blockIdx.x];
    const int m = mt.at_mol[i];
    const int a0 = mt.mol_atom[m];
    const int nao = mt.mol_nao[m];
    const int spi = mt.at_sp[i];
    const int nci = sp.ncart[spi];
    const double xi = A.xyz[3 * i], yi = A.xyz[3 * i + 1], zi = A.xyz[3 * i + 2];
    const long long dense0 = mt.mol_dense[m];
    const int cart_m = mt.mol_cart[m];
    const int ao_m = mt.mol_ao[m];
    const int iao0 = mt.at_ao[i];
    const int rows_cap = TW_GRAD / nci;
    // half list: j < i, plus j == i for lattice self-images; the adjoint also
    // visits the on-site block (j == i, T = 0) for dE/dse.
    const bool with_self = (mt.ntrans > 1) || (MODE == GRAD_ADJOINT);
    const int jend = with_self ? i + 1 : i;
    const int cend = (jend > a0) ? (mt.at_cart[jend - 1] + sp.ncart[mt.at_sp[jend - 1]]) : 0;
    for (int k = threadIdx.x; k < 3 + MAXSH + 9; k += NTHREADS) accI[k] = 0.0;
    if (threadIdx.x == 0) s_count = 0;

    int c0 = 0;
    while (c0 < cend) {
        const int jfirst = mt.cart_atom[cart_m + c0];
        int lo = jfirst, hi = jend;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            int end = mt.at_cart[mid] + sp.ncart[mt.at_sp[mid]];
            if (end <= c0 + rows_cap) lo = mid; else hi = mid;
        }
        const int jlast = lo;
        const int c1 = mt.at_cart[jlast] + sp.ncart[mt.at_sp[jlast]];
        const int ntask = (c1 - c0) * nci;
        const int nwin = jlast - jfirst + 1;
        __syncthreads();
        for (int k = threadIdx.x; k < nwin * (3 + MAXSH); k += NTHREADS)
            accJ[k / (3 + MAXSH)][k % (3 + MAXSH)] = 0.0;
        __syncthreads();

        for (int t = threadIdx.x; t < ntask; t += NTHREADS) {
            const int cjl = t / nci, ci = t - cjl * nci;
            const int j = mt.cart_atom[cart_m + c0 + cjl];
            const int spj = mt.at_sp[j];
            const int cj = c0 + cjl - mt.at_cart[j];
            const int slj = sp.cart_sh[spj * TBX_MAXCART + cj];
            const int sli = sp.cart_sh[spi * TBX_MAXCART + ci];
            const int expj = sp.cart_exp[spj * TBX_MAXCART + cj];
            const int expi = sp.cart_exp[spi * TBX_MAXCART + ci];
            int ej[3], ei[3];
            unpack_exp(expj, ej);
            unpack_exp(expi, ei);
            const int kj = expj >> 12, ki = expi >> 12;  // index within the shell
            const int slotj = slj + sp.msh * spj, sloti = sli + sp.msh * spi;
            const int lj = sp.ang[slotj], li = sp.ang[sloti];

            // Cartesian weights from the spherical ones (transpose of the transform)
            double wS = 0.0, wH = 0.0, wHon = 0.0, wDi[3] = {0, 0, 0}, wDj[3] = {0, 0, 0};
            double wQi[6] = {0, 0, 0, 0, 0, 0}, wQj[6] = {0, 0, 0, 0, 0, 0};
            const int muj0 = mt.at_ao[j] + sp.sh_ao[slotj];
            const int nui0 = iao0 + sp.sh_ao[sloti];
            for (int a = 0; a < kCartN[lj][kj]; ++a) {
                const int mu = muj0 + kCartM[lj][kj][a];
                for (int b = 0; b < kCartN[li][ki]; ++b) {
                    const int nu = nui0 + kCartM[li][ki][b];
                    const double w = kCartW[lj][kj][a] * kCartW[li][ki][b];
                    const long long e1 = dense0 + (long long)nu * nao + mu;  // (mu, nu)
                    const long long e2 = dense0 + (long long)mu * nao + nu;  // (nu, mu)
                    if (MODE == GRAD_PLAIN) {
                        const double p = A.P[e1];
                        double vs = 0.0;
                        for (int s = 0; s < A.nspin; ++s)
                            vs += A.P[e1 + s * A.pstride] *
                                  (A.vao[ao_m + mu + s * A.vstride] + A.vao[ao_m + nu + s * A.vstride]);
                        wS += w * (-vs - 2.0 * A.W[e1]);
                        wH += w * 2.0 * p;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            wDi[k] -= w * p * A.vdp[3 * i + k];
                            wDj[k] -= w * p * A.vdp[3 * j + k];
                        }
#pragma unroll
                        for (int k = 0; k < 6; ++k) {
                            wQi[k] -= w * p * A.vqp[6 * i + k];
                            wQj[k] -= w * p * A.vqp[6 * j + k];
                        }
                    } else {
                        wS += w * (A.Sb[e1] + A.Sb[e2]);
                        wH += w * (A.Hb[e1] + A.Hb[e2]);
                        wHon += w * A.Hb[e1];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            wDi[k] += w * A.Db[3 * e1 + k];
                            wDj[k] += w * A.Db[3 * e2 + k];
                        }
#pragma unroll
                        for (int k = 0; k < 6; ++k) {
                            wQi[k] += w * A.Qb[6 * e1 + k];
                            wQj[k] += w * A.Qb[6 * e2 + k];
                        }
                    }
                }
            }
            // quadrupoles are contracted after trace removal: fold the
            // (symmetric) traceless map into the weights
            {
                const double ti = 0.5 * (wQi[0] + wQi[2] + wQi[5]);
                const double tj = 0.5 * (wQj[0] + wQj[2] + wQj[5]);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    wQi[k] *= 1.5;
                    wQj[k] *= 1.5;
                }
                wQi[0] -= ti; wQi[2] -= ti; wQi[5] -= ti;
                wQj[0] -= tj; wQj[2] -= tj; wQj[5] -= tj;
            }

            const int npj = sp.nprim[slotj], npi = sp.nprim[sloti];
            const double dx0 = xi - A.xyz[3 * j], dy0 = yi - A.xyz[3 * j + 1],
                         dz0 = zi - A.xyz[3 * j + 2];
            const double hsc = sp.hscale[slj + sp.msh * (sli + sp.msh * (spj + sp.nsp * spi))];
            const double kpi = sp.shpoly[sloti], kpj = sp.shpoly[slotj];
            const double radij = sp.rad[spi] + sp.rad[spj];
            const double sesum = 0.5 * (A.se[mt.at_sh[i] + sli] + A.se[mt.at_sh[j] + slj]);
            double dG[3] = {0, 0, 0}, dsi = 0.0, dsj = 0.0;
            double sig[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            int nimg = 0;
            const bool self = (j == i);
            for (int itr = 0; itr < mt.ntrans; ++itr) {
                const double vx = dx0 - mt.trans[3 * itr], vy = dy0 - mt.trans[3 * itr + 1],
                             vz = dz0 - mt.trans[3 * itr + 2];
                const double r2 = vx * vx + vy * vy + vz * vz;
                const double vec[3] = {vx, vy, vz};
                if (r2 < kEps) {
                    if (MODE == GRAD_ADJOINT && self) {
                        // on-site block: only dE/dse, H = S (se_i + se_j)/2
                        double s = 0.0;
                        for (int ip = 0; ip < npi; ++ip) {
                            const double ai = sp.alpha[sloti * MAXG + ip];
                            const double ci_ = sp.coeff[sloti * MAXG + ip];
                            for (int jp = 0; jp < npj; ++jp) {
                                const double aj = sp.alpha[slotj * MAXG + jp];
                                const double oab = 1.0 / (ai + aj);
                                const double so = sqrt(oab);
                                const double cc = kSqrtPi3 * so * so * so * ci_ *
                                                  sp.coeff[slotj * MAXG + jp];
                                const Mom M = make_mom(0.5 * oab);
                                double V[3][3];
#pragma unroll
                                for (int k = 0; k < 3; ++k) fwd1d(ei[k], ej[k], 0.0, 0.0, M, V[k]);
                                s += cc * V[0][0] * V[1][0] * V[2][0];
                            }
                        }
                        dsi += 0.5 * wHon * s;
                        dsj += 0.5 * wHon * s;
                    }
                    continue;
                }
                if (r2 > sp.cutoff2) continue;
                ++nimg;
                const double rr = sqrt(sqrt(r2) / radij);
                const double shi = 1.0 + kpi * rr, shj = 1.0 + kpj * rr;
                const double shp = shi * shj;
                const double dshp = (shi * kpj + shj * kpi) * 0.5 * rr / r2;
                const double hs = hsc * shp;
                const double hij = sesum * hs;
                double a3[3] = {0, 0, 0}, b3[3] = {0, 0, 0}, s = 0.0;
                for (int ip = 0; ip < npi; ++ip) {
                    const double ai = sp.alpha[sloti * MAXG + ip];
                    const double ci_ = sp.coeff[sloti * MAXG + ip];
                    for (int jp = 0; jp < npj; ++jp) {
                        const double aj = sp.alpha[slotj * MAXG + jp];
                        const double eab = ai + aj;
                        const double oab = 1.0 / eab;
                        const double est = ai * aj * r2 * oab;
                        if (est > sp.intcut) continue;
                        const double so = sqrt(oab);
                        const double cc = exp(-est) * kSqrtPi3 * so * so * so * ci_ *
                                          sp.coeff[slotj * MAXG + jp];
                        const Mom M = make_mom(0.5 * oab);
                        double Vi[3][3], Vj[3][3], Gi[3][3], Gj[3][3];
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            grad1d(ei[k], ej[k], -vec[k] * aj * oab, vec[k] * ai * oab, ai, M,
                                   Vi[k], Vj[k], Gi[k], Gj[k]);
                        s += cc * wH * (Vi[0][0] * Vi[1][0] * Vi[2][0]);
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            // f_k(m): direction d carries the derivative
                            double fi[3][3], fj[3][3];
#pragma unroll
                            for (int k = 0; k < 3; ++k)
#pragma unroll
                                for (int mm = 0; mm < 3; ++mm) {
                                    fi[k][mm] = (k == d) ? Gi[k][mm] : Vi[k][mm];
                                    fj[k][mm] = (k == d) ? Gj[k][mm] : Vj[k][mm];
                                }
                            const double ds = fi[0][0] * fi[1][0] * fi[2][0];
                            double acc = wS * ds;
                            acc += wDi[0] * fi[0][1] * fi[1][0] * fi[2][0];
                            acc += wDi[1] * fi[0][0] * fi[1][1] * fi[2][0];
                            acc += wDi[2] * fi[0][0] * fi[1][0] * fi[2][1];
                            acc += wDj[0] * fj[0][1] * fj[1][0] * fj[2][0];
                            acc += wDj[1] * fj[0][0] * fj[1][1] * fj[2][0];
                            acc += wDj[2] * fj[0][0] * fj[1][0] * fj[2][1];
#pragma unroll
                            for (int q = 0; q < 6; ++q) {
                                acc += wQi[q] * fi[0][kQm[q][0]] * fi[1][kQm[q][1]] * fi[2][kQm[q][2]];
                                acc += wQj[q] * fj[0][kQm[q][0]] * fj[1][kQm[q][1]] * fj[2][kQm[q][2]];
                            }
                            a3[d] += cc * acc;
                            b3[d] += cc * wH * ds;
                        }
                    }
                }
                double g[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) g[d] = a3[d] + hij * (b3[d] + s * dshp / shp * vec[d]);
                const double fself = self ? 0.5 : 1.0;
                dsi += fself * 0.5 * hs * s;
                dsj += fself * 0.5 * hs * s;
                const double fs = self ? 0.25 : 0.5;
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) sig[a + 3 * b] += fs * (vec[a] * g[b] + g[a] * vec[b]);
                if (!self) {
                    dG[0] += g[0];
                    dG[1] += g[1];
                    dG[2] += g[2];
                }
            }
            if (A.nnl && ci == 0 && cj == 0 && nimg > 0) atomicAdd(&s_count, nimg);
            // scatter to LDS accumulators
            const int jw = j - jfirst;
            if (dG[0] != 0.0 || dG[1] != 0.0 || dG[2] != 0.0) {
                atomicAdd(&accJ[jw][0], -dG[0]);
                atomicAdd(&accJ[jw][1], -dG[1]);
                atomicAdd(&accJ[jw][2], -dG[2]);
                atomicAdd(&accI[0], dG[0]);
                atomicAdd(&accI[1], dG[1]);
                atomicAdd(&accI[2], dG[2]);
            }
            if (dsj != 0.0) atomicAdd(&accJ[jw][3 + slj], dsj);
            if (dsi != 0.0) atomicAdd(&accI[3 + sli], dsi);
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (sig[k] != 0.0) atomicAdd(&accI[3 + MAXSH + k], sig[k]);
        }
        __syncthreads();
        // flush window atoms to global
        for (int k = threadIdx.x; k < nwin; k += NTHREADS) {
            const int j = jfirst + k;
            if (j == i) {
                // self-images / on-site: merge into atom i's accumulators
                for (int q = 0; q < 3 + MAXSH; ++q) atomicAdd(&accI[q], accJ[k][q]);
                continue;
            }
            for (int q = 0; q < 3; ++q)
                if (accJ[k][q] != 0.0) unsafeAtomicAdd(&A.grad[3 * j + q], accJ[k][q]);
            const int nshj = sp.nsh[mt.at_sp[j]];
            for (int q = 0; q < nshj; ++q)
                if (accJ[k][3 + q] != 0.0) unsafeAtomicAdd(&A.dse[mt.at_sh[j] + q], accJ[k][3 + q]);
        }
        c0 = c1;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        if (accI[threadIdx.x] != 0.0) unsafeAtomicAdd(&A.grad[3 * i + threadIdx.x], accI[threadIdx.x]);
    } else if (threadIdx.x < 3 + MAXSH) {
        const int q = threadIdx.x - 3;
        if (q < sp.nsh[spi] && accI[threadIdx.x] != 0.0)
            unsafeAtomicAdd(&A.dse[mt.at_sh[i] + q], accI[threadIdx.x]);
    } else if (threadIdx.x < 3 + MAXSH + 9) {
        const int q = threadIdx.x - 3 - MAXSH;
        if (accI[threadIdx.x] != 0.0) unsafeAtomicAdd(&A.sigma[9 * m + q], accI[threadIdx.x]);
    }
    if (A.nnl && threadIdx.x == 0 && s_count != A.nnl[i]) atomicOr(A.flag, 1);
}

// dEdcn (plain mode): sum over shells of dsedcn * dE/dse plus the on-site
// term dsedcn * sum diag(P) of the shell (xtb/h0.f90:654-663).
struct CnArgs {
    TbxSpecies sp;
    TbxMols mol;
    const double *dse, *dsedcn, *P;
    double *dEdcn;
};

__global__ void k_dedcn(CnArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.mol.nat) return;
    const int m = A.mol.at_mol[i];
    const int spi = A.mol.at_sp[i];
    const int nao = A.mol.mol_nao[m];
    const long long dense0 = A.mol.mol_dense[m];
    double acc = 0.0;
    for (int ish = 0; ish < A.sp.nsh[spi]; ++ish) {
        const int sh = A.mol.at_sh[i] + ish;
        const int slot = ish + A.sp.msh * spi;
        double diag = 0.0;
        const int mu0 = A.mol.at_ao[i] + A.sp.sh_ao[slot];
        for (int k = 0; k < 2 * A.sp.ang[slot] + 1; ++k)
            diag += A.P[dense0 + (long long)(mu0 + k) * nao + mu0 + k];
        acc += A.dsedcn[sh] * (A.dse[sh] + diag);
    }
    A.dEdcn[i] = acc;
}

// ---------------------------------------------------------------------------
// launchers (host side of this translation unit)

extern "C" hipError_t tbx_launch_forward(const TbxSpecies *sp, const TbxMols *mol,
                                         const double *xyz, const double *se, const int *order,
                                         double *S, double *D, double *Q, double *H,
                                         const int *nnl, int *flag, hipStream_t stream)
{
    FwdArgs a{*sp, *mol, xyz, se, order, S, D, Q, H, nnl, flag};
    (void)hipGetLastError();  // launch errors are read below; drop stale ones
    hipLaunchKernelGGL(k_forward, dim3(mol->nat), dim3(NTHREADS), 0, stream, a);
    return hipGetLastError();
}

extern "C" hipError_t tbx_launch_gradient(int mode, const TbxSpecies *sp, const TbxMols *mol,
                                          const double *xyz, const double *se, const int *order,
                                          const double *P, const double *W, const double *vao,
                                          const double *vdp, const double *vqp, int nspin,
                                          long long pstride, long long vstride,
                                          const double *Sb, const double *Db, const double *Qb,
                                          const double *Hb, double *grad, double *dse,
                                          double *sigma, const int *nnl, int *flag,
                                          hipStream_t stream)
{
    GradArgs a{*sp, *mol, xyz, se, order, P, W, vao, vdp, vqp, nspin, pstride, vstride,
               Sb, Db, Qb, Hb, grad, dse, sigma, nnl, flag};
    (void)hipGetLastError();
    if (mode == GRAD_PLAIN)
        hipLaunchKernelGGL(k_gradient<GRAD_PLAIN>, dim3(mol->nat), dim3(NTHREADS), 0, stream, a);
    else
        hipLaunchKernelGGL(k_gradient<GRAD_ADJOINT>, dim3(mol->nat), dim3(NTHREADS), 0, stream, a);
    return hipGetLastError();
}

extern "C" hipError_t tbx_launch_dedcn(const TbxSpecies *sp, const TbxMols *mol,
                                       const double *dse, const double *dsedcn, const double *P,
                                       double *dEdcn, hipStream_t stream)
{
    CnArgs a{*sp, *mol, dse, dsedcn, P, dEdcn};
    (void)hipGetLastError();
    const int nb = (mol->nat + 255) / 256;
    hipLaunchKernelGGL(k_dedcn, dim3(nb), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// Fill kernel (used instead of hipMemsetAsync so that callers may capture the
// launches in a HIP graph).
__global__ void k_fill(double *p, long long n, double v)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

extern "C" hipError_t tbx_launch_fill(double *p, long long n, double v, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const long long nb = (n + 255) / 256;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}
