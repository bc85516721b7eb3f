// tbx_kernels.hip -- gfx950 kernels of the GFN2-xTB integral / H0 path.
//
// Work units.  The host cuts every molecule's half list (atom pairs i, j <= i
// and lattice images) into work units of consecutive pairs, at most one
// workgroup's worth each (tbx_tables.h: TbxUnitHead and its plan), and
// packs into each unit's plan everything that does not depend on the
// geometry: pair records, per-segment rows, the unit's shell pairs (type-
// major) and the per-type task offsets.  One workgroup runs one unit; it
// copies the plan to LDS with one 16-byte load per lane.  With a geometry
// hint the pairs of a molecule closer than cutoff + skin come first and the
// others fill units of their own (TBX_UNIT_FAR); k_hint_check tells every
// launch whether any atom moved more than skin/2 from the hint, and while
// none did those units are known to hold no pair inside the cutoff.
//
// Forward, k_forward (get_hamiltonian, xtb/h0.f90:220-429): a unit's lanes
// form the pair vectors, then one lane per shell pair counts its kept
// primitive pairs (the sorted primitive-pair tables make them a prefix) and
// its H0 factor; one lane per Cartesian-pair task schedules its chunks of
// kept primitive pairs, a block scan packs the chunks into work items, and
// each item accumulates the 10 Cartesian moments (S, D, Q about centre i) of
// its task in LDS (on-site blocks copy the species' same-centre moments).
// Then one lane per dense element applies the Cartesian -> spherical map
// (integral/trafo.f90), removes the quadrupole trace and stores the block
// (rows j, column i, operator on i) or its mirror (rows i, column j,
// operator shifted to j, h0.f90:704-747): every element of a molecule is
// written exactly once, so molecules need no zero-fill.  Far units only
// store their zero blocks.  k_forward<true> builds H0 from the overlap
// scaled in the diatomic frame of each pair (integral/diat_trafo.f90).
//
// Gradient and adjoint, k_gradient<MODE> (get_hamiltonian_gradient,
// h0.f90:432-666, and the backward of the forward): one lane per task gathers
// the spherical weights of its block (P, W and the potentials, or the
// adjoints of S, H0, D, Q) into a Cartesian task record, with the j-centred
// moment weights shifted to centre i; work items recompute the primitive
// integrals and their R_i derivatives (Obara-Saika tables shared by the
// Cartesian pairs of s/p shell pairs) and contract them with the record at
// once -- no derivative tensor is stored.  Given the forward's S and D of a
// molecule batch, the terms through the moment shift and the H0 distance
// polynomial are contracted per element from them instead
// (explicit_element), and the records shrink to the 10 moment weights.  Item results are summed across
// the wavefront with DPP segmented scans, then per pair in LDS, and leave as
// FP64 atomics into dE/dR, dE/dse and sigma.  Far units are not dispatched
// while the hint holds; k_gradient_far runs them (persistent) when it does
// not.  k_onsite adds the on-site dE/dse of the adjoint.
//
// All arithmetic is FP64.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>

#include "tbx_tables.h"

// Workgroup sizes: one lane per Cartesian-pair task of a work unit
#define NTHREADS TBX_UNIT_TASKS        // forward
#define GTHREADS TBX_GRAD_UNIT_TASKS   // gradient
#define MAXSH 4

static constexpr double kS3 = 1.7320508075688772935;
static constexpr double kSqrtPi3 = 5.5683279968317078453;  // pi^(3/2)
static constexpr double kEps = 2.220446049250313e-16;

// Cartesian -> spherical: sph(l, m) = sum_k w * cart(l, c_k), CCA Cartesian
// order; l=1 (y, z, x); l=2 m=-2..2: sqrt3 xy, sqrt3 yz, zz-(xx+yy)/2,
// sqrt3 xz, sqrt3/2 (xx-yy).
static constexpr int kSphN[3][5] = {{1, 0, 0, 0, 0}, {1, 1, 1, 0, 0}, {1, 1, 3, 1, 2}};
static constexpr int kSphC[3][5][3] = {{{0}},
                                   {{1}, {2}, {0}},
                                   {{1}, {4}, {5, 0, 3}, {2}, {0, 3}}};
static constexpr double kSphW[3][5][3] = {
    {{1.0}},
    {{1.0}, {1.0}, {1.0}},
    {{kS3}, {kS3}, {1.0, -0.5, -0.5}, {kS3}, {0.5 * kS3, -0.5 * kS3}}};
// Transpose of the same map: Cartesian c -> (m, w) pairs.
static constexpr int kCartN[3][6] = {{1, 0, 0, 0, 0, 0}, {1, 1, 1, 0, 0, 0}, {2, 1, 1, 2, 1, 1}};
static constexpr int kCartM[3][6][2] = {{{0}},
                                    {{2}, {0}, {1}},
                                    {{2, 4}, {0}, {3}, {2, 4}, {1}, {2}}};
static constexpr double kCartW[3][6][2] = {
    {{1.0}},
    {{1.0}, {1.0}, {1.0}},
    {{-0.5, 0.5 * kS3}, {kS3}, {kS3}, {-0.5, -0.5 * kS3}, {kS3}, {1.0}}};


// The same maps as arithmetic on packed constants (no memory access), for
// use with lane-dependent l, m and Cartesian indices in the kernels.
// sph(l, m) = sum_{a < sph_n} sph_w * cart(l, sph_c)
__host__ __device__ __forceinline__ int sph_n(int l, int m)
{
    return (l == 2) ? (m == 2 ? 3 : (m == 4 ? 2 : 1)) : 1;
}
__host__ __device__ __forceinline__ int sph_c(int l, int m, int a)
{
    if (l == 0) return 0;
    if (l == 1) return (0x021 >> (4 * m)) & 15;
    return (int)((0x30002305004001ULL >> (4 * (3 * m + a))) & 15);
}
__host__ __device__ __forceinline__ double sph_w(int l, int m, int a)
{
    if (l < 2) return 1.0;
    if (m == 2) return a == 0 ? 1.0 : -0.5;
    if (m == 4) return a == 0 ? 0.5 * kS3 : -0.5 * kS3;
    return kS3;
}
// transpose: cart(l, c) enters sph(l, cart_m) with weight cart_w
__host__ __device__ __forceinline__ int cart_n(int l, int c)
{
    return (l == 2 && (c == 0 || c == 3)) ? 2 : 1;
}
__host__ __device__ __forceinline__ int cart_m(int l, int c, int k)
{
    if (l == 0) return 0;
    if (l == 1) return (0x102 >> (4 * c)) & 15;
    return (int)((0x20142030042ULL >> (4 * (2 * c + k))) & 15);
}
__host__ __device__ __forceinline__ double cart_w(int l, int c, int k)
{
    if (l < 2 || c == 5) return 1.0;
    if (c == 0) return k == 0 ? -0.5 : 0.5 * kS3;
    if (c == 3) return k == 0 ? -0.5 : -0.5 * kS3;
    return kS3;
}
// exponent of direction k (x, y, z) in Cartesian function c (CCA order) of shell l
__host__ __device__ __forceinline__ int cart_exp(int l, int c, int k)
{
    if (l == 0) return 0;
    if (l == 1) return c == k ? 1 : 0;
    return (int)((0x814211142ULL >> (6 * c + 2 * k)) & 3);
}

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

// integral of sum_k P[k] y^k (y + c)^m, m = 0, 1, 2; N coefficients.  Only
// the even moments are nonzero, so each k takes the terms of its parity
// alone (a product with a literal 0.0 is not folded under IEEE rules and
// would still be issued).
template <int N>
__device__ __forceinline__ void mom3(const double *P, double c, const Mom &M, double &i0,
                                     double &i1, double &i2)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if ((k & 1) == 0) {  // y^k: moments k (m = 0, 1, 2) and k + 2 (m = 2)
            const double mk = mom(M, k), mk2 = mom(M, k + 2);
            a0 += P[k] * mk;
            a1 += P[k] * (c * mk);
            a2 += P[k] * (mk2 + c * c * mk);
        } else {             // moment k + 1 (m = 1, 2)
            const double mk1 = mom(M, k + 1);
            a1 += P[k] * mk1;
            a2 += P[k] * (2.0 * c * mk1);
        }
    }
    i0 = a0;
    i1 = a1;
    i2 = a2;
}

// quadrupole component q -> moment exponent per direction (xx xy yy xz yz zz)
static constexpr int kQm[6][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};

// ---------------------------------------------------------------------------
// Gradient / adjoint contraction.

enum { GRAD_PLAIN = 0, GRAD_ADJOINT = 1 };

struct GradArgs {
    TbxSpecies sp;
    TbxMols mol;
    const double *xyz, *se;
    // plain: P, W dense per molecule (spin planes at +spin*pstride), vao [nao_tot*nspin],
    // vdp [3*nat*nspin], vqp [6*nat*nspin]
    const double *P, *W, *vao, *vdp, *vqp;
    int nspin;
    long long pstride;   // elements between spin planes of P (and W)
    long long vstride;   // elements between spin planes of vao
    // adjoint: dense adjoints of S, D, Q, H0
    const double *Sb, *Db, *Qb, *Hb;
    // the forward's S and D (operator on the column atom) at the same
    // geometry, molecules only: the explicit pair-vector terms are then
    // contracted once per element (explicit_element) instead of in the work
    // items; NULL otherwise
    const double *Sf, *Df;
    // outputs (accumulated with atomics; caller zeroes)
    double *grad;   // [3*nat]
    double *dse;    // [nsh]  dE/d(selfenergy)
    double *sigma;  // [9*nmol]
    const int *nnl; // optional [nat] list check
    int *flag;
    int chunk;      // primitive pairs per work item
};

// Cartesian exponents (CCA order) of the functions of a shell, l <= 2.
static constexpr int kCartExp[3][6][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

__device__ __forceinline__ constexpr int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// c[0..N-1] = coefficients of (y + p)^n, runtime n in [0, N-1].
template <int N>
__device__ __forceinline__ void pow_poly(int n, double p, double *c)
{
    c[0] = 1.0;
#pragma unroll
    for (int k = 1; k < N; ++k) c[k] = 0.0;
#pragma unroll
    for (int r = 0; r < N - 1; ++r) {
        const bool on = r < n;
#pragma unroll
        for (int k = N - 1; k >= 1; --k) c[k] = on ? c[k - 1] + p * c[k] : c[k];
        c[0] = on ? p * c[0] : c[0];
    }
}

// Gradient 1D factors with compile-time degree bounds (a <= LI, b <= LJ),
// moments about the centre A of function i:
//   V[m] = int (y+pa)^(a+m) (y+pb)^b,   G[m] = d/dA_x V[m]
// (function i and the moment operator both move with A).
template <int LI, int LJ>
__device__ __forceinline__ void grad1d_t(int a, int b, double pa, double pb, double ai,
                                         const Mom &M, double V[3], double G[3])
{
    double A0[LI + 1], Am[LI + 1], B[LJ + 1], A1[LI + 2], Ga[LI + 2];
    double P[LI + LJ + 1], PG[LI + LJ + 2];
    pow_poly<LI + 1>(a, pa, A0);
    pow_poly<LI + 1>(a > 0 ? a - 1 : 0, pa, Am);
    pow_poly<LJ + 1>(b, pb, B);
    A1[0] = pa * A0[0];
#pragma unroll
    for (int k = 1; k <= LI; ++k) A1[k] = A0[k - 1] + pa * A0[k];
    A1[LI + 1] = A0[LI];
    const double ta = 2.0 * ai, fa = (double)a;
#pragma unroll
    for (int k = 0; k <= LI; ++k) Ga[k] = ta * A1[k] - fa * Am[k];
    Ga[LI + 1] = ta * A1[LI + 1];
#pragma unroll
    for (int k = 0; k < LI + LJ + 1; ++k) P[k] = 0.0;
#pragma unroll
    for (int k = 0; k < LI + LJ + 2; ++k) PG[k] = 0.0;
#pragma unroll
    for (int x = 0; x <= LI; ++x)
#pragma unroll
        for (int y = 0; y <= LJ; ++y) P[x + y] += A0[x] * B[y];
#pragma unroll
    for (int x = 0; x <= LI + 1; ++x)
#pragma unroll
        for (int y = 0; y <= LJ; ++y) PG[x + y] += Ga[x] * B[y];
    mom3<LI + LJ + 1>(P, pa, M, V[0], V[1], V[2]);
    double g1, g2;
    mom3<LI + LJ + 2>(PG, pa, M, G[0], g1, g2);
    G[1] = g1 - V[0];
    G[2] = g2 - 2.0 * V[1];
}

// sum over the moments (S, D, Q about A) of w * F_x F_y F_z
__device__ __forceinline__ double contract_moments(const double *fx, const double *fy,
                                                   const double *fz, double wS,
                                                   const double wD[3], const double wQ[6])
{
    // pair the y and z factors first
    const double yz00 = fy[0] * fz[0], yz10 = fy[1] * fz[0], yz01 = fy[0] * fz[1];
    const double yz20 = fy[2] * fz[0], yz11 = fy[1] * fz[1], yz02 = fy[0] * fz[2];
    const double t0 = wS * yz00 + wD[1] * yz10 + wD[2] * yz01 + wQ[2] * yz20 + wQ[4] * yz11 +
                      wQ[5] * yz02;
    const double t1 = wD[0] * yz00 + wQ[1] * yz10 + wQ[3] * yz01;
    const double t2 = wQ[0] * yz00;
    return fx[0] * t0 + fx[1] * t1 + fx[2] * t2;
}

// The three derivative contractions of one Cartesian pair: c[k] is
// contract_moments with the derivative factors G in direction k and the
// moment factors V in the other two.  c[1] and c[2] share the x factors V[0],
// so the weights are folded with them once: u_ab = sum over the x moments of
// w V0, the weight of fy[a] fz[b].
__device__ __forceinline__ void contract_grad(const double G[3][3], const double V[3][3],
                                              const double w[10], double c[3])
{
    const double *wD = w + 1, *wQ = w + 4;  // D: x y z; Q: xx xy yy xz yz zz
    c[0] = contract_moments(G[0], V[1], V[2], w[0], wD, wQ);
    const double u00 = w[0] * V[0][0] + wD[0] * V[0][1] + wQ[0] * V[0][2];
    const double u10 = wD[1] * V[0][0] + wQ[1] * V[0][1];
    const double u01 = wD[2] * V[0][0] + wQ[3] * V[0][1];
    const double u20 = wQ[2] * V[0][0], u11 = wQ[4] * V[0][0], u02 = wQ[5] * V[0][0];
    c[1] = V[2][0] * (u00 * G[1][0] + u10 * G[1][1] + u20 * G[1][2]) +
           V[2][1] * (u01 * G[1][0] + u11 * G[1][1]) + V[2][2] * (u02 * G[1][0]);
    c[2] = G[2][0] * (u00 * V[1][0] + u10 * V[1][1] + u20 * V[1][2]) +
           G[2][1] * (u01 * V[1][0] + u11 * V[1][1]) + G[2][2] * (u02 * V[1][0]);
}

// packed symmetric index (xx, xy, yy, xz, yz, zz) of the 3x3 element a + 3b
static constexpr int kSigPack[9] = {0, 1, 3, 1, 2, 4, 3, 4, 5};

// LDS accumulators of the row atom i of a gradient workgroup (one copy per
// wavefront): dE/dR_i (3), dE/dse of its shells (MAXSH), sigma packed (6)
#define ACCI_N (3 + MAXSH + 6)
// primitive pairs per work item when the caller passes none (runtime-tunable per launch)
#define FWD_CHUNK_DEFAULT 3
#define GRAD_CHUNK_DEFAULT 1
#define NREC 20    // doubles of a task record

// Task record fields (LDS, structure of arrays over the tasks of a unit)
enum {
    R_W = 0,     // 10 contraction weights about centre i: S, D (3), Q (6)
    R_EX = 10,   // weights of S in the explicit v-derivative (3)
    R_WQ = 13,   // traceless-folded Q weights of the j-centred moments (6)
    R_HSW = 19,  // dE/dse_i = dE/dse_j per unit S
};

// Atom pair of a unit, staged in LDS
using PairInfo = TbxPair;

// shell, m and l of spherical function ao of a species with shell layout desc
__device__ __forceinline__ void ao_decode(int desc, int ao, int &sh, int &m, int &l)
{
    const int nsh = desc & 7;
    sh = 0;
    m = ao;
    bool done = false;
#pragma unroll
    for (int s = 0; s < MAXSH - 1; ++s) {
        const int n = 2 * ((desc >> (3 + 2 * s)) & 3) + 1;
        const bool adv = !done && (s < nsh - 1) && (m >= n);
        if (adv) {
            m -= n;
            sh = s + 1;
        } else {
            done = true;
        }
    }
    l = (desc >> (3 + 2 * sh)) & 3;
}

// first spherical AO of shell sh of a species with shell layout desc
__device__ __forceinline__ int ao_first(int desc, int sh)
{
    int n = 0;
#pragma unroll
    for (int s = 0; s < MAXSH - 1; ++s)
        if (s < sh) n += 2 * ((desc >> (3 + 2 * s)) & 3) + 1;
    return n;
}

// inclusive prefix sum over the 64 lanes of a wavefront
__device__ __forceinline__ int wave_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(v, d, 64);
        if (lane >= d) v += y;
    }
    return v;
}

// Number of primitive pairs of a slot pair kept by the screening at r2: the
// list is sorted by mu, so the kept pairs are its leading entries.
__device__ __forceinline__ int kept_pairs(const double *mu, int n, double r2, double intcut)
{
    if (n > 64) {
        int lo = 0, hi = n;  // invariant: entries < lo kept, entries >= hi dropped
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (mu[mid] * r2 <= intcut) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    // two rounds of eight independent loads: the blocks of eight entries
    // fully kept, then the kept entries of the first block that is not
    int b = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int q = 8 * k + 7;
        b += (q < n && mu[q] * r2 <= intcut) ? 1 : 0;
    }
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int q = 8 * b + k;
        c += (q < n && mu[q] * r2 <= intcut) ? 1 : 0;
    }
    return 8 * b + c;
}

// last index k in [0, n) with off[k] <= c (off non-decreasing, off[0] <= c)
__device__ __forceinline__ int find_last_le(const int *off, int n, int c)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= c) lo = mid; else hi = mid;
    }
    return lo;
}

// Work item -> task: last index k in [0, NT) with off[k] <= c (off
// non-decreasing, off[0] = 0 <= c), in two rounds of independent LDS reads
// (the block of 16 entries, then the entry within it) instead of eight
// dependent ones.  off must be 16-byte aligned.
template <int NT>
__device__ __forceinline__ int find_task(const int *off, int c)
{
    static_assert(NT % 16 == 0 && NT <= 256, "two-level search over blocks of 16");
    int b = 0;
#pragma unroll
    for (int k = 1; k < NT / 16; ++k) b += (off[16 * k] <= c) ? 1 : 0;
    const int4 *blk = reinterpret_cast<const int4 *>(off + 16 * b);
    int t = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int4 v = blk[k];
        t += (v.x <= c ? 1 : 0) + (v.y <= c ? 1 : 0) + (v.z <= c ? 1 : 0) + (v.w <= c ? 1 : 0);
    }
    return 16 * b + t - 1;
}

// Unit of workgroup b of n: XCD-contiguous ranges (tbx_xcd_block), so that
// neighbouring units of a molecule (the same dense columns and cache lines)
// share one L2.
__device__ __forceinline__ int xcd_unit(int b, int n) { return tbx_xcd_block(b, n); }

// Unit staging shared by the forward and gradient kernels (LDS): the unit's
// plan (TbxUnitHead) and the pair vectors, for units of at most NT tasks
// (= workgroup size) and NP atom pairs.
template <int NT, int NP>
struct __align__(16) UnitStageT {
    static constexpr int kThreads = NT;
    PairInfo pinf[NP];
    int pext[NP];                   // lattice vector << 5 | segment
    int4 sinfo[TBX_UNIT_SEGS];      // first pair, first row AO, rows | mirrored rows << 16,
                                    // AOs of atom i
    int spl[NT];                    // shell pairs: pair | jsh << 8 | ish << 12 | type << 16
                                    // (type-major)
    int toff[10], tsp[10];          // task / shell-pair offsets per type
    double pv[NP][3];               // v = R_i - R_j - T
    int wsum[NT / 64][9];
};
using FwdStage = UnitStageT<TBX_UNIT_TASKS, TBX_UNIT_PAIRS>;
using GradStage = UnitStageT<TBX_GRAD_UNIT_TASKS, TBX_GRAD_UNIT_PAIRS>;
static_assert(sizeof(PairInfo) == 40, "pair record layout");
#define TBX_STAGE_LAYOUT(US)                                                                   \
    static_assert(offsetof(US, pext) % 16 == 0 && offsetof(US, sinfo) % 16 == 0 &&             \
                      offsetof(US, spl) % 16 == 0 && offsetof(US, toff) % 16 == 0 &&           \
                      offsetof(US, tsp) == offsetof(US, toff) + 40,                            \
                  "plan sections are copied as 16-byte words")
TBX_STAGE_LAYOUT(FwdStage);
TBX_STAGE_LAYOUT(GradStage);
#undef TBX_STAGE_LAYOUT

// (pair, jsh, ish, type) of a packed shell-pair entry
__device__ __forceinline__ int4 unpack_spl(int v)
{
    return make_int4(v & 255, (v >> 8) & 15, (v >> 12) & 15, v >> 16);
}

// Number of 16-byte words of a unit's plan staged in LDS
__device__ __forceinline__ int plan_words(const TbxUnitHead &H)
{
    const int npair = H.counts & 255, nseg = (H.counts >> 8) & 255, nspair = H.counts >> 16;
    return ((10 * npair + 3) >> 2) + ((npair + 3) >> 2) + nseg + ((nspair + 3) >> 2) + 5;
}

// Copy a unit's plan (layout: TbxUnitHead) into LDS, one 16-byte word per
// lane and step.
template <class US>
__device__ __forceinline__ void load_unit(const TbxUnitHead &H, const int *plan, US &U)
{
    const int npair = H.counts & 255, nseg = (H.counts >> 8) & 255, nspair = H.counts >> 16;
    const int4 *src = reinterpret_cast<const int4 *>(plan) + H.plan;
    const int e1 = (10 * npair + 3) >> 2;
    const int e2 = e1 + ((npair + 3) >> 2);
    const int e3 = e2 + nseg;
    const int e4 = e3 + ((nspair + 3) >> 2);
    const int total = e4 + 5;
    int4 *base = reinterpret_cast<int4 *>(&U);
    for (int k = threadIdx.x; k < total; k += US::kThreads) {
        const int4 v = src[k];
        int w;  // destination word within U
        if (k < e1) w = k;
        else if (k < e2) w = (int)(offsetof(US, pext) / 16) + (k - e1);
        else if (k < e3) w = (int)(offsetof(US, sinfo) / 16) + (k - e2);
        else if (k < e4) w = (int)(offsetof(US, spl) / 16) + (k - e3);
        else w = (int)(offsetof(US, toff) / 16) + (k - e4);
        base[w] = v;
    }
}

// v = R_i - R_j - T of a pair
__device__ __forceinline__ void pair_vec(const TbxMols &mt, const double *xyz, const PairInfo &pr,
                                         int pext, double v[3])
{
    const int itr = pext >> 5;
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = xyz[3 * pr.i + c] - xyz[3 * pr.j + c] - mt.trans[3 * itr + c];
}

// Decode task tid of a staged unit: shell pair index, type and Cartesian pair.
template <class US>
__device__ __forceinline__ int4 decode_task(const US &U, int task, int &spx)
{
    int t = 0;
    while (t < 8 && task >= U.toff[t + 1]) ++t;
    const int ncj = ncart_of(t % 3);
    const int ncp = ncart_of(t / 3) * ncj;
    const int local = task - U.toff[t];
    spx = U.tsp[t] + local / ncp;
    const int cp = local % ncp;
    return make_int4(t, cp / ncj, cp % ncj, cp);  // (type, ki, kj, cp)
}

// Shell-pair phase of a unit (one lane per shell pair): screening count and
// first entry of the sorted primitive-pair list, and the H0 factors
// hs = hscale * shpoly(R), dsr = (d shpoly/dr) / (r shpoly), sesum =
// (se_i + se_j)/2.  onsite: the j == i, T = 0 block counts (forward only).
template <class US>
__device__ __forceinline__ bool shellpair_setup(const TbxSpecies &sp, const TbxMols &mt,
                                                const double *xyz, const US &U,
                                                const double *se, int k, bool onsite_ok,
                                                int &kept, int &pi0, double &hs, double &sesum,
                                                double &dsr)
{
    const int4 e = unpack_spl(U.spl[k]);
    const PairInfo pr = U.pinf[e.x];
    const int ish = e.z, jsh = e.y;
    const double v[3] = {U.pv[e.x][0], U.pv[e.x][1], U.pv[e.x][2]};
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const int sloti = ish + sp.msh * pr.spi, slotj = jsh + sp.msh * pr.spj;
    pi0 = sp.pp_off[sloti * sp.nslot + slotj];
    kept = 0;
    hs = sesum = dsr = 0.0;
    const bool onsite = onsite_ok && (pr.i == pr.j) && r2 < kEps;
    if (!onsite && (r2 < kEps || r2 > sp.cutoff2)) return false;
    const int npp = sp.pp_off[sloti * sp.nslot + slotj + 1] - pi0;
    kept = kept_pairs(sp.pp_mu + pi0, npp, onsite ? 0.0 : r2, sp.intcut);
    sesum = 0.5 * (se[pr.shi + ish] + se[pr.shj + jsh]);
    if (onsite) {
        hs = 1.0;  // on-site block: H = S (se_i + se_j)/2
        return true;
    }
    const double hsc = sp.hscale[jsh + sp.msh * (ish + sp.msh * (pr.spj + sp.nsp * pr.spi))];
    const double kpi = sp.shpoly[sloti], kpj = sp.shpoly[slotj];
    const double rr = sqrt(sqrt(r2) / (sp.rad[pr.spi] + sp.rad[pr.spj]));
    const double shi = 1.0 + kpi * rr, shj = 1.0 + kpj * rr;
    const double shp = shi * shj;
    hs = hsc * shp;
    dsr = (shi * kpj + shj * kpi) * 0.5 * rr / (r2 * shp);
    return false;
}

// block-wide exclusive scan of one int per lane; returns the total
template <int NT>
__device__ __forceinline__ int block_scan(int v, int *out, int (*wsum)[9])
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int inc = wave_scan(v);
    __syncthreads();  // wsum may still be read by the caller's previous phase
    if (lane == 63) wsum[wv][0] = inc;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += wsum[w][0];
    out[tid] = base + inc - v;
    if (tid == NT - 1) out[NT] = base + inc;
    __syncthreads();
    return out[NT];
}


// ---------------------------------------------------------------------------
// Forward: dense S, D, Q, H0 over work units of the half list (j <= i).

struct FwdArgs {
    TbxSpecies sp;
    TbxMols mol;
    const double *xyz;   // [3*nat]
    const double *se;    // [nsh]
    double *S, *D, *Q, *H;
    int accumulate;      // 1: add into zeroed outputs (lattice images), 0: plain stores
    int chunk;           // primitive pairs per work item
    const int *stale;    // nonzero: the pair-order hint no longer holds (far units run whole)
};

// 1D factors V[m] = int (y+pa)^(a+m) (y+pb)^b, m = 0..2 (moments about A)
template <int LI, int LJ>
__device__ __forceinline__ void fwd1d_t(int a, int b, double pa, double pb, const Mom &M,
                                        double V[3])
{
    double A0[LI + 1], B[LJ + 1], P[LI + LJ + 1];
    pow_poly<LI + 1>(a, pa, A0);
    pow_poly<LJ + 1>(b, pb, B);
#pragma unroll
    for (int k = 0; k < LI + LJ + 1; ++k) P[k] = 0.0;
#pragma unroll
    for (int x = 0; x <= LI; ++x)
#pragma unroll
        for (int y = 0; y <= LJ; ++y) P[x + y] += A0[x] * B[y];
    mom3<LI + LJ + 1>(P, pa, M, V[0], V[1], V[2]);
}

// Primitive pairs [p0, p1) of one Cartesian pair: S, D (3), Q (6, raw) about i
template <int LI, int LJ>
__device__ __forceinline__ void fchunk_run(const double *__restrict__ pp, const double vec[3],
                                           int ki, int kj, int p0, int p1, double o[10])
{
    const int ei[3] = {cart_exp(LI, ki, 0), cart_exp(LI, ki, 1), cart_exp(LI, ki, 2)};
    const int ej[3] = {cart_exp(LJ, kj, 0), cart_exp(LJ, kj, 1), cart_exp(LJ, kj, 2)};
    const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
    // entries are loaded one iteration ahead
    double nx[TBX_PP_N];
#pragma unroll
    for (int f = 0; f < TBX_PP_N; ++f) nx[f] = pp[TBX_PP_N * p0 + f];
    for (int p = p0; p < p1; ++p) {
        const double mu = nx[TBX_PP_MU], ajo = nx[TBX_PP_AJO], aio = nx[TBX_PP_AIO];
        const double h = nx[TBX_PP_H], pre = nx[TBX_PP_PRE];
        if (p + 1 < p1) {
#pragma unroll
            for (int f = 0; f < TBX_PP_N; ++f) nx[f] = pp[TBX_PP_N * (p + 1) + f];
        }
        const double cc = exp(-(mu * r2)) * pre;
        const Mom M = make_mom(h);
        double V[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) fwd1d_t<LI, LJ>(ei[k], ej[k], -vec[k] * ajo, vec[k] * aio, M, V[k]);
        // the x factors scaled by cc and the six y-z products, each formed once
        const double cx0 = cc * V[0][0], cx1 = cc * V[0][1], cx2 = cc * V[0][2];
        const double yz00 = V[1][0] * V[2][0], yz10 = V[1][1] * V[2][0], yz01 = V[1][0] * V[2][1];
        const double yz20 = V[1][2] * V[2][0], yz11 = V[1][1] * V[2][1], yz02 = V[1][0] * V[2][2];
        o[0] += cx0 * yz00;
        o[1] += cx1 * yz00;
        o[2] += cx0 * yz10;
        o[3] += cx0 * yz01;
        o[4] += cx2 * yz00;  // quadrupole xx xy yy xz yz zz (kQm)
        o[5] += cx1 * yz10;
        o[6] += cx0 * yz20;
        o[7] += cx1 * yz01;
        o[8] += cx0 * yz11;
        o[9] += cx0 * yz02;
    }
}

__device__ __forceinline__ void put(double *p, double v, bool acc)
{
    if (acc) unsafeAtomicAdd(p, v); else *p = v;
}

// A forward unit none of whose pairs is within the cutoff: its blocks are
// zero.  They are the segments' rectangles, block A (rows of the atoms j,
// columns of atom i) and the mirror (rows of atom i, columns of the atoms
// j < i), so only the segment words of the plan are needed (read here unless
// `staged`: a far unit, TBX_UNIT_FAR, known from the pair-order hint while
// every atom stays within skin/2 of it, stages nothing else).  off: LDS of
// 2 * TBX_UNIT_SEGS + 1 ints.
__device__ __forceinline__ void far_zero_unit(const FwdArgs &A, const TbxUnitHead &H, FwdStage &U,
                                              int *off, bool staged)
{
    const int tid = threadIdx.x;
    const int npair = H.counts & 255, nseg = (H.counts >> 8) & 255;
    if (!staged) {
        const int4 *src = reinterpret_cast<const int4 *>(A.mol.fplan) + H.plan +
                          ((10 * npair + 3) >> 2) + ((npair + 3) >> 2);
        if (tid < nseg) U.sinfo[tid] = src[tid];
    }
    __syncthreads();
    if (tid == 0) tbx_elem_offsets(U.sinfo, nseg, off);  // block A and mirror of each segment
    __syncthreads();
    const int total = off[2 * nseg];
    const long long dense0 = H.dense0;
    const int nao = H.nao;
    const double2 z = make_double2(0.0, 0.0);
    for (int e = tid; e < total; e += NTHREADS) {
        const int k2 = find_last_le(off, 2 * nseg, e);
        const int4 g = U.sinfo[k2 >> 1];  // aoi, aoj0, rows | mrows << 16, naoi
        const int local = e - off[k2];
        long long idx;
        if ((k2 & 1) == 0) {  // block A, column-major: consecutive rows
            const int rows = g.z & 0xffff;
            const int c = local / rows;
            idx = dense0 + (long long)(g.x + c) * nao + g.y + (local - c * rows);
        } else {  // mirror: rows of atom i, runs of naoi
            const int c = local / g.w;
            idx = dense0 + (long long)(g.y + c) * nao + g.x + (local - c * g.w);
        }
        A.S[idx] = 0.0;
        A.H[idx] = 0.0;
        const bool odd = (idx & 1) != 0;
        *reinterpret_cast<double2 *>(A.D + 3 * idx + (odd ? 1 : 0)) = z;
        A.D[3 * idx + (odd ? 0 : 2)] = 0.0;
        double2 *q2 = reinterpret_cast<double2 *>(A.Q + 6 * idx);
        q2[0] = z;
        q2[1] = z;
        q2[2] = z;
    }
}

// --- diatomic-frame scaling of the overlap in H0 (integral/diat_trafo.f90)
// The bond vector's polar angles with the reference's conventions (harmtr,
// :241-262): on the z axis phi is taken as 0, in the xy plane sin(theta) = 1.
struct DiatFrame {
    double ct, st, cp, sp;
};
__device__ __forceinline__ DiatFrame diat_frame(const double v[3])
{
    const double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double n0 = v[0] / len, n1 = v[1] / len, ct = v[2] / len;
    DiatFrame F;
    F.ct = ct;
    if (fabs(ct) == 1.0) {
        F.st = 0.0;
        F.cp = 1.0;
        F.sp = 0.0;
    } else if (fabs(ct) == 0.0) {
        F.st = 1.0;
        F.cp = n0;
        F.sp = n1;
    } else {
        F.st = sqrt(n0 * n0 + n1 * n1);
        F.cp = n0 / F.st;
        F.sp = n1 / F.st;
    }
    return F;
}

// Rotation of shell l to the diatomic frame (harmtr, :264-335): entry (r, c),
// r a function of the molecular frame, c one of the bond frame (both 0..2l,
// p in y, z, x order, d in m = -2..2 order); the bond axis is c = l.
__device__ double diat_rot(const DiatFrame &F, int l, int r, int c)
{
    if (l == 0) return 1.0;
    const double ct = F.ct, st = F.st, cp = F.cp, sp = F.sp;
    if (l == 1) {
        switch (r + 3 * c) {
        case 0: return cp;
        case 1: return 0.0;
        case 2: return -sp;
        case 3: return st * sp;
        case 4: return ct;
        case 5: return st * cp;
        case 6: return ct * sp;
        case 7: return -st;
        default: return ct * cp;
        }
    }
    const double c2t = ct * ct - st * st, s2t = 2.0 * st * ct;
    const double c2p = cp * cp - sp * sp, s2p = 2.0 * sp * cp;
    switch (r + 5 * c) {
    case 0: return ct * c2p;
    case 1: return -st * cp;
    case 2: return 0.0;
    case 3: return st * sp;
    case 4: return -ct * s2p;
    case 5: return st * c2p;
    case 6: return ct * cp;
    case 7: return 0.0;
    case 8: return -ct * sp;
    case 9: return -st * s2p;
    case 10: return kS3 * st * st * s2p * 0.5;
    case 11: return kS3 * s2t * sp * 0.5;
    case 12: return (3.0 * ct * ct - 1.0) * 0.5;
    case 13: return kS3 * s2t * cp * 0.5;
    case 14: return kS3 * st * st * c2p * 0.5;
    case 15: return s2t * s2p * 0.5;
    case 16: return c2t * sp;
    case 17: return -kS3 * s2t * 0.5;
    case 18: return c2t * cp;
    case 19: return s2t * c2p * 0.5;
    case 20: return (1.0 + ct * ct) * s2p * 0.5;
    case 21: return -s2t * sp * 0.5;
    case 22: return kS3 * st * st * 0.5;
    case 23: return -s2t * cp * 0.5;
    default: return (1.0 + ct * ct) * c2p * 0.5;
    }
}

// Element (mj, mi) of the scaled overlap of a shell pair (lj on atom j, li on
// atom i) at bond vector v: in the bond frame T = R_j^T S R_i, the elements
// of equal m (m = -L..L, L = min(lj, li), at T[lj + m][li + m]) are scaled by
// ksig, kpi or kdel for |m| = 0, 1, 2 (scale_diatomic_frame, :642-689), and
// the block is taken back, R_j T' R_i^T (diat_trafo, :99-156).  The others
// are unchanged and R is orthogonal, so the element is sv plus the rotated
// changes.  S(c, d) of the shell pair comes from its Cartesian tasks
// (task0 first, row-major by the Cartesian function of atom i).
__device__ double diat_element(const double (*acc)[10], int task0, int ncj, int lj, int li,
                               int mj, int mi, const double v[3], const double *k3, double sv)
{
    if (!(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 0.0)) return sv;
    const DiatFrame F = diat_frame(v);
    const int L = min(lj, li);
    double T[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // T[m + 2]
    for (int c = 0; c < 2 * lj + 1; ++c)
        for (int d = 0; d < 2 * li + 1; ++d) {
            double sc = 0.0;
            for (int a = 0; a < sph_n(lj, c); ++a)
                for (int b = 0; b < sph_n(li, d); ++b)
                    sc += sph_w(lj, c, a) * sph_w(li, d, b) *
                          acc[task0 + sph_c(li, d, b) * ncj + sph_c(lj, c, a)][0];
#pragma unroll
            for (int m = -2; m <= 2; ++m)
                if (m >= -L && m <= L)
                    T[m + 2] += diat_rot(F, lj, c, lj + m) * sc * diat_rot(F, li, d, li + m);
        }
    double out = sv;
#pragma unroll
    for (int m = -2; m <= 2; ++m)
        if (m >= -L && m <= L)
            out += (k3[m < 0 ? -m : m] - 1.0) * T[m + 2] * diat_rot(F, lj, mj, lj + m) *
                   diat_rot(F, li, mi, li + m);
    return out;
}

// One workgroup per forward unit.  (1) stage pairs and shell pairs; (2) one
// lane per Cartesian-pair task counts its kept primitive pairs, the first
// lane of each shell pair also forms its H0 factor; (3) chunks of kept pairs
// are scanned into work items; (4) work items accumulate Cartesian S, D, Q
// about i in LDS; (5) one lane per dense element applies the spherical
// transform and writes the block (rows j, column i: operator on i) and its
// mirror (rows i, column j: operator shifted to j, xtb/h0.f90:704-747) --
// every element of the molecule's matrices is written exactly once.  DIAT:
// off-site H0 blocks from the diatomic-frame scaled overlap (sp.kdiat).
template <bool DIAT>
__global__ void __launch_bounds__(NTHREADS, 2) k_forward(FwdArgs A)
{
    __shared__ FwdStage U;
    __shared__ __align__(16) double acc[TBX_UNIT_TASKS][10];  // S, D (3), Q (6) per task
    __shared__ double hfac[TBX_UNIT_TASKS];  // H0 / S per shell pair
    __shared__ int spk[TBX_UNIT_TASKS], spp[TBX_UNIT_TASKS];  // kept pairs, first entry
    __shared__ int tinfo[TBX_UNIT_TASKS];    // type | ki << 4 | kj << 8 | shell pair << 12 | pair << 20
    __shared__ __align__(16) int coff[TBX_UNIT_TASKS + 1];  // chunk offsets, then element offsets
    __shared__ unsigned char spmap[TBX_UNIT_PAIRS * 16];     // first task of (pair, ish, jsh)
    __shared__ unsigned char rowpair[2 * TBX_UNIT_TASKS];    // pair of each block-A row
    const TbxSpecies &sp = A.sp;
    const TbxMols &mt = A.mol;
    const int tid = threadIdx.x;
    const TbxUnitHead H = mt.funit[xcd_unit(blockIdx.x, gridDim.x)];
    const bool accum = A.accumulate != 0;
    if ((H.nel & TBX_UNIT_FAR) && *A.stale == 0) {
        // outputs with lattice images are zero-filled before the launch
        if (!accum) far_zero_unit(A, H, U, coff, false);
        return;
    }
    const int npair = H.counts & 255, nseg = (H.counts >> 8) & 255, nspair = H.counts >> 16;
    const int nao = H.nao;
    const long long dense0 = H.dense0;
    const int chunk = A.chunk;
    for (int k = tid; k < 10 * TBX_UNIT_TASKS; k += NTHREADS) (&acc[0][0])[k] = 0.0;
    // (1) the unit's plan
    load_unit(H, mt.fplan, U);
    const int nel = H.nel & TBX_NEL_MASK;  // <= 2 * TBX_UNIT_TASKS: two elements per lane
    __syncthreads();

    // (2) pair vectors.  A unit with no pair inside the cutoff (on-site
    // pairs are inside) has all-zero blocks: it writes its zeros and ends
    // (the host groups each molecule's distant pairs into units of their own
    // when given a geometry hint).
    bool live = false;
    if (tid < npair) {
        double v[3];
        pair_vec(mt, A.xyz, U.pinf[tid], U.pext[tid], v);
#pragma unroll
        for (int c = 0; c < 3; ++c) U.pv[tid][c] = v[c];
        live = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] <= sp.cutoff2;
    }
    if (!__syncthreads_or(live)) {
        if (!accum) far_zero_unit(A, H, U, coff, true);
        return;
    }
    // shell pairs: screening counts and H0 factors (one lane per shell pair)
    for (int k = tid; k < nspair; k += NTHREADS) {
        int kept, pi0;
        double hs, sesum, dsr;
        const bool on = shellpair_setup(sp, mt, A.xyz, U, A.se, k, true, kept, pi0, hs, sesum, dsr);
        spk[k] = on ? -1 : kept;  // on-site blocks: moments from the species table
        spp[k] = pi0;
        hfac[k] = sesum * hs;
    }
    __syncthreads();
    // tasks: one Cartesian pair per lane
    const int ntask = U.toff[9];
    int nch = 0;
    if (tid < ntask) {
        int spx;
        const int4 d = decode_task(U, tid, spx);
        tinfo[tid] = d.x | (d.y << 4) | (d.z << 8) | (spx << 12) | ((U.spl[spx] & 255) << 20);
        if (d.w == 0) {  // the shell pair's first task, for the element phase
            const int4 e = unpack_spl(U.spl[spx]);  // pair, jsh, ish, type
            spmap[16 * e.x + 4 * e.z + e.y] = (unsigned char)tid;
        }
        const int kept = spk[spx];
        if (kept >= 0) {
            nch = (kept + chunk - 1) / chunk;
        } else {
            // on-site block (j == i, no lattice shift): the same-centre moments
            // of the species, no work items
            const int4 e = unpack_spl(U.spl[spx]);  // pair, jsh, ish, type
            const int spi = U.pinf[e.x].spi;
            const int ci = sp.sh_cart[e.z + sp.msh * spi] + d.y;
            const int cj = sp.sh_cart[e.y + sp.msh * spi] + d.z;
            const double2 *src = reinterpret_cast<const double2 *>(
                sp.onsite + ((size_t)(spi * TBX_MAXCART + ci) * TBX_MAXCART + cj) * 10);
            double2 *dst = reinterpret_cast<double2 *>(acc[tid]);
#pragma unroll
            for (int q = 0; q < 5; ++q) dst[q] = src[q];
        }
    }
    // (3) chunk offsets
    const int nchunk = block_scan<NTHREADS>(nch, coff, U.wsum);

    // (4) work items
    for (int c = tid; c < nchunk; c += NTHREADS) {
        const int tk = find_last_le(coff, NTHREADS, c);
        const int info = tinfo[tk];
        const int t = info & 15, ki = (info >> 4) & 15, kj = (info >> 8) & 15;
        const int spx = (info >> 12) & 255;
        const int p = info >> 20;
        const int q = c - coff[tk];
        const int p0 = spp[spx] + q * chunk;
        const int p1 = spp[spx] + min(spk[spx], (q + 1) * chunk);
        const double vec[3] = {U.pv[p][0], U.pv[p][1], U.pv[p][2]};
        double o[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        switch (t) {
        case 0: fchunk_run<0, 0>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 1: fchunk_run<0, 1>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 2: fchunk_run<0, 2>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 3: fchunk_run<1, 0>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 4: fchunk_run<1, 1>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 5: fchunk_run<1, 2>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 6: fchunk_run<2, 0>(sp.pp, vec, ki, kj, p0, p1, o); break;
        case 7: fchunk_run<2, 1>(sp.pp, vec, ki, kj, p0, p1, o); break;
        default: fchunk_run<2, 2>(sp.pp, vec, ki, kj, p0, p1, o); break;
        }
#pragma unroll
        for (int k = 0; k < 10; ++k) atomicAdd(&acc[tk][k], o[k]);
    }

    // (5) one lane per dense element, located from the segment words
    // (tbx_elem_locate): block A (rows of atoms j, the columns of atom i) and
    // the mirror (rows of atom i, columns of the j < i); every element of the
    // molecule's matrices is written by exactly one lane of one unit
    __syncthreads();  // the items are done with coff
    int *roff = coff + 2 * nseg + 1;  // row offsets after the element offsets
    if (tid == 0) {
        tbx_elem_offsets(U.sinfo, nseg, coff);
        tbx_row_offsets(U.sinfo, nseg, roff);
    }
    __syncthreads();
    if (tid < npair) tbx_row_map_pair(U.sinfo, roff, U.pinf, U.pext, npair, tid, rowpair);
    __syncthreads();
    for (int base = 0; base < nel; base += NTHREADS) {
        const int e = base + tid;
        if (e < nel) {
            const TbxElem el = tbx_elem_locate_map(coff, nseg, U.sinfo, U.pinf, roff, rowpair, e);
            const int p = el.p, nul = el.nul, muj = el.muj;
            const bool mirror = el.mirror != 0;
            const PairInfo pr = U.pinf[p];
            int ish, mi, li, jsh, mj, lj;
            ao_decode(pr.di, nul, ish, mi, li);
            ao_decode(pr.dj, muj, jsh, mj, lj);
            const int task0 = spmap[16 * p + 4 * ish + jsh];
            const int spx = (tinfo[task0] >> 12) & 255;
            const int ncj = ncart_of(lj);
            const int mu = pr.aoj + muj;
            double v[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) v[q] = 0.0;
            const int nsj = sph_n(lj, mj), nsi = sph_n(li, mi);
            for (int a = 0; a < nsj; ++a) {
                const double wa = sph_w(lj, mj, a);
                const int cj = sph_c(lj, mj, a);
                for (int b = 0; b < nsi; ++b) {
                    const double w = wa * sph_w(li, mi, b);
                    const int tk = task0 + sph_c(li, mi, b) * ncj + cj;
                    const double2 *a2 = reinterpret_cast<const double2 *>(acc[tk]);
#pragma unroll
                    for (int q = 0; q < 5; ++q) {
                        const double2 x = a2[q];
                        v[2 * q] += w * x.x;
                        v[2 * q + 1] += w * x.y;
                    }
                }
            }
            const double sv = v[0];
            const double tr = 0.5 * (v[4] + v[6] + v[9]);
            double d[3] = {v[1], v[2], v[3]};
            double qv[6] = {1.5 * v[4] - tr, 1.5 * v[5], 1.5 * v[6] - tr,
                            1.5 * v[7], 1.5 * v[8], 1.5 * v[9] - tr};
            const int nu = pr.aoi + nul;
            long long idx;
            if (!mirror) {
                idx = dense0 + (long long)nu * nao + mu;
            } else {
                // operator on j: shift the moments from i to j (x - R_j = (x - R_i) + v)
                idx = dense0 + (long long)mu * nao + nu;
                const double vv[3] = {U.pv[p][0], U.pv[p][1], U.pv[p][2]};
                double sq[6];
                sq[0] = 2 * vv[0] * d[0] + vv[0] * vv[0] * sv;
                sq[2] = 2 * vv[1] * d[1] + vv[1] * vv[1] * sv;
                sq[5] = 2 * vv[2] * d[2] + vv[2] * vv[2] * sv;
                sq[1] = vv[0] * d[1] + vv[1] * d[0] + vv[0] * vv[1] * sv;
                sq[3] = vv[0] * d[2] + vv[2] * d[0] + vv[0] * vv[2] * sv;
                sq[4] = vv[1] * d[2] + vv[2] * d[1] + vv[1] * vv[2] * sv;
                const double ts = 0.5 * (sq[0] + sq[2] + sq[5]);
                qv[0] += 1.5 * sq[0] - ts;
                qv[1] += 1.5 * sq[1];
                qv[2] += 1.5 * sq[2] - ts;
                qv[3] += 1.5 * sq[3];
                qv[4] += 1.5 * sq[4];
                qv[5] += 1.5 * sq[5] - ts;
                for (int c = 0; c < 3; ++c) d[c] += vv[c] * sv;
            }
            double hv = sv;  // overlap entering H0
            if (DIAT && spk[spx] >= 0) {
                const double vv[3] = {U.pv[p][0], U.pv[p][1], U.pv[p][2]};
                hv = diat_element(acc, task0, ncj, lj, li, mj, mi, vv,
                                  sp.kdiat + 3 * (pr.spi + sp.nsp * pr.spj), sv);
            }
            if (accum) {
                put(A.S + idx, sv, true);
                put(A.H + idx, hfac[spx] * hv, true);
#pragma unroll
                for (int c = 0; c < 3; ++c) put(A.D + 3 * idx + c, d[c], true);
#pragma unroll
                for (int c = 0; c < 6; ++c) put(A.Q + 6 * idx + c, qv[c], true);
            } else {
                A.S[idx] = sv;
                A.H[idx] = hfac[spx] * hv;
                // the 3 dipole components: one 16-byte store (aligned for either
                // parity of idx) and one 8-byte store
                const bool odd = (idx & 1) != 0;
                *reinterpret_cast<double2 *>(A.D + 3 * idx + (odd ? 1 : 0)) =
                    odd ? make_double2(d[1], d[2]) : make_double2(d[0], d[1]);
                A.D[3 * idx + (odd ? 0 : 2)] = odd ? d[0] : d[2];
                // the 6 quadrupole components of an element start on a 16-byte boundary
                double2 *q2 = reinterpret_cast<double2 *>(A.Q + 6 * idx);
                q2[0] = make_double2(qv[0], qv[1]);
                q2[1] = make_double2(qv[2], qv[3]);
                q2[2] = make_double2(qv[4], qv[5]);
            }
        }
    }
}


// The explicit pair-vector terms of one spherical element (mu on atom j, nu
// on atom i; e1 = (mu, nu), e2 = (nu, mu)) from the forward's S and
// i-centred D there.  Per Cartesian pair the work items would add
// S_c ex_c + 2 Wj_c D_c and S_c hsw_c (task_setup's R_EX, R_WQ, R_HSW), with
// weights that are the spherical weights X pulled back through the
// spherical transform T; since sum_c (T^T X T)_c S_c = sum_e X_e (T S T^T)_e,
// the same sum runs over the block's elements with the forward's spherical
// S and D.  explicit_core takes the element's weights: hsum (of H0), db (the
// j-centred dipole), wq (the j-centred quadrupole, trace map folded in), and
// adds its share of the pair sum dE/dv to g and of dE/dse_i = dE/dse_j to ds.
__device__ __forceinline__ void explicit_core(const GradArgs &A, long long e1, int i, int j,
                                              const double v[3], double hs, double sesum,
                                              double dsr, double hsum, const double db[3],
                                              const double wq[6], double g[3], double &ds)
{
    const double s = A.Sf[e1];
    const double d[3] = {A.Df[3 * e1], A.Df[3 * e1 + 1], A.Df[3 * e1 + 2]};
    const double hd = sesum * hs * hsum * dsr;
    const double X[3] = {db[0] + 2.0 * (wq[0] * v[0] + 0.5 * (wq[1] * v[1] + wq[3] * v[2])) + hd * v[0],
                         db[1] + 2.0 * (wq[2] * v[1] + 0.5 * (wq[1] * v[0] + wq[4] * v[2])) + hd * v[1],
                         db[2] + 2.0 * (wq[5] * v[2] + 0.5 * (wq[3] * v[0] + wq[4] * v[1])) + hd * v[2]};
    g[0] += s * X[0] + 2.0 * (wq[0] * d[0] + 0.5 * (wq[1] * d[1] + wq[3] * d[2]));
    g[1] += s * X[1] + 2.0 * (wq[2] * d[1] + 0.5 * (wq[1] * d[0] + wq[4] * d[2]));
    g[2] += s * X[2] + 2.0 * (wq[5] * d[2] + 0.5 * (wq[3] * d[0] + wq[4] * d[1]));
    ds += (j == i ? 0.5 : 1.0) * 0.5 * hs * hsum * s;
}

// explicit_core for an element whose weights are read here (shell pairs with
// a d shell, where Cartesian and spherical elements are not one to one)
template <int MODE>
__device__ __forceinline__ void explicit_element(const GradArgs &A, long long e1, long long e2,
                                                 int i, int j, const double v[3], double hs,
                                                 double sesum, double dsr, double g[3],
                                                 double &ds)
{
    double hsum, db[3], qb[6];
    if (MODE == GRAD_PLAIN) {  // weights -P x (atom j's potentials), as in task_setup
        const double p = A.P[e1];
        hsum = 2.0 * p;
#pragma unroll
        for (int k = 0; k < 3; ++k) db[k] = -p * A.vdp[3 * j + k];
#pragma unroll
        for (int k = 0; k < 6; ++k) qb[k] = -p * A.vqp[6 * j + k];
    } else {
        hsum = A.Hb[e1] + A.Hb[e2];
#pragma unroll
        for (int k = 0; k < 3; ++k) db[k] = A.Db[3 * e2 + k];
#pragma unroll
        for (int k = 0; k < 6; ++k) qb[k] = A.Qb[6 * e2 + k];
    }
    const double tq = 0.5 * (qb[0] + qb[2] + qb[5]);  // the trace-removal map
    double wq[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) wq[k] = 1.5 * qb[k];
    wq[0] -= tq;
    wq[2] -= tq;
    wq[5] -= tq;
    explicit_core(A, e1, i, j, v, hs, sesum, dsr, hsum, db, wq, g, ds);
}

// Task set-up: one Cartesian function pair (ki of shell ish on atom i, kj of
// shell jsh on atom j) at lattice vector itr.  Folds the spherical weights of
// the pair's block into Cartesian ones (transpose of the spherical map),
// shifts the j-centred multipole weights to centre i and stores the record.
// Returns the number of primitive pairs kept by the screening.
// Cyclic axis frames of the p-p row items (pprow_run): the frame of row k
// has axis s along (k + s) mod 3.  rot3 picks x_k; the vector and the packed
// symmetric tensor (xx xy yy xz yz zz) in that frame are rot_vec, rot_sym.
__device__ __forceinline__ double rot3(double x0, double x1, double x2, int k)
{
    return k == 0 ? x0 : (k == 1 ? x1 : x2);
}
__device__ __forceinline__ void rot_vec(const double x[3], int k, double y[3])
{
    y[0] = rot3(x[0], x[1], x[2], k);
    y[1] = rot3(x[1], x[2], x[0], k);
    y[2] = rot3(x[2], x[0], x[1], k);
}
__device__ __forceinline__ void rot_sym(const double q[6], int k, double y[6])
{
    y[0] = rot3(q[0], q[2], q[5], k);
    y[1] = rot3(q[1], q[4], q[3], k);
    y[2] = rot3(q[2], q[5], q[0], k);
    y[3] = rot3(q[3], q[1], q[4], k);
    y[4] = rot3(q[4], q[3], q[1], k);
    y[5] = rot3(q[5], q[0], q[2], k);
}

template <int MODE, bool EX>
__device__ void task_setup(const GradArgs &A, const PairInfo &pr, const double v[3], int ish,
                           int jsh, int li, int lj, int ki, int kj, int nao, long long dense0,
                           int ao_m, double hs, double sesum, double dsr, double *rec, int ld,
                           double gx[3], double &dsx)
{
    const int i = pr.i, j = pr.j;
    const double vx = v[0], vy = v[1], vz = v[2];
    const int muj0 = pr.aoj + ao_first(pr.dj, jsh);
    const int nui0 = pr.aoi + ao_first(pr.di, ish);
    const int na = cart_n(lj, kj), nb = cart_n(li, ki);
    double wS = 0.0, wH = 0.0;
    double wDi[3] = {0, 0, 0}, wDj[3] = {0, 0, 0};
    double wQi[6] = {0, 0, 0, 0, 0, 0}, wQj[6] = {0, 0, 0, 0, 0, 0};
    if (MODE == GRAD_PLAIN) {
        double pt = 0.0;
        for (int a = 0; a < na; ++a) {
            const int mu = muj0 + cart_m(lj, kj, a);
            for (int b = 0; b < nb; ++b) {
                const int nu = nui0 + cart_m(li, ki, b);
                const double w = cart_w(lj, kj, a) * cart_w(li, ki, b);
                const long long e1 = dense0 + (long long)nu * nao + mu;  // (mu, nu)
                double vs = 0.0;
                for (int sgn = 0; sgn < A.nspin; ++sgn)
                    vs += A.P[e1 + sgn * A.pstride] *
                          (A.vao[ao_m + mu + sgn * A.vstride] + A.vao[ao_m + nu + sgn * A.vstride]);
                wS += w * (-vs - 2.0 * A.W[e1]);
                pt += w * A.P[e1];
            }
        }
        // the multipole-potential weights factorise: -P x (per-atom vector)
        wH = 2.0 * pt;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            wDi[k] = -pt * A.vdp[3 * i + k];
            wDj[k] = -pt * A.vdp[3 * j + k];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            wQi[k] = -pt * A.vqp[6 * i + k];
            wQj[k] = -pt * A.vqp[6 * j + k];
        }
    } else {
        for (int a = 0; a < na; ++a) {
            const int mu = muj0 + cart_m(lj, kj, a);
            for (int b = 0; b < nb; ++b) {
                const int nu = nui0 + cart_m(li, ki, b);
                const double w = cart_w(lj, kj, a) * cart_w(li, ki, b);
                const long long e1 = dense0 + (long long)nu * nao + mu;  // (mu, nu)
                const long long e2 = dense0 + (long long)mu * nao + nu;  // (nu, mu)
                wS += w * (A.Sb[e1] + A.Sb[e2]);
                wH += w * (A.Hb[e1] + A.Hb[e2]);
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
    // quadrupoles are contracted after trace removal: fold the map into the weights
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
    // Moments about j are shifted moments about i (x - B = (x - A) + v, v = A - B):
    //   D_j = D_i + v S,   Q_j = Q_i + v D_i^T + D_i v^T + v v^T S,
    // so the j-centred weights fold into i-centred ones plus an explicit v
    // dependence, and only i-centred 1D factors are needed.
    const double Wv[3] = {wQj[0] * vx + 0.5 * (wQj[1] * vy + wQj[3] * vz),
                          wQj[2] * vy + 0.5 * (wQj[1] * vx + wQj[4] * vz),
                          wQj[5] * vz + 0.5 * (wQj[3] * vx + wQj[4] * vy)};
    // H0 = S * (se_i + se_j)/2 * hscale * shpoly(R): hs = hscale * shpoly and
    // dsr = (d shpoly / d r) / (r shpoly) come from the shell-pair phase
    const double hw = sesum * hs * wH;
    // p-p shell pairs run by rows (pprow_run) in the frame of the row: the
    // record of Cartesian pair (ki, kj) goes to slot (kj - ki) mod 3 of its
    // row, with its vector and tensor weights in that frame
    const int fr = (li == 1 && lj == 1) ? ki : 0;
    double *rw = (li == 1 && lj == 1) ? rec + ((kj - ki + 3) % 3 - kj) : rec;
    rw[(R_W + 0) * ld] = wS + hw + wDj[0] * vx + wDj[1] * vy + wDj[2] * vz + Wv[0] * vx +
                         Wv[1] * vy + Wv[2] * vz;
    {
        const double wd[3] = {wDi[0] + wDj[0] + 2.0 * Wv[0], wDi[1] + wDj[1] + 2.0 * Wv[1],
                              wDi[2] + wDj[2] + 2.0 * Wv[2]};
        double wq[6], y[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) wq[k] = wQi[k] + wQj[k];
        rot_vec(wd, fr, y);
#pragma unroll
        for (int k = 0; k < 3; ++k) rw[(R_W + 1 + k) * ld] = y[k];
        rot_sym(wq, fr, y);
#pragma unroll
        for (int k = 0; k < 6; ++k) rw[(R_W + 4 + k) * ld] = y[k];
    }
    if (!EX) {
        // s/p shell pairs: the Cartesian pair is one spherical element, whose
        // weights are the ones gathered above; its explicit terms (the lane
        // owns it) go to gx, dsx.  Shell pairs with a d shell: explicit_element.
        if (li <= 1 && lj <= 1) {
            const long long e1 = dense0 + (long long)(nui0 + cart_m(li, ki, 0)) * nao + muj0 +
                                 cart_m(lj, kj, 0);
            explicit_core(A, e1, i, j, v, hs, sesum, dsr, wH, wDj, wQj, gx, dsx);
        }
        return;
    }
    const double hd = hw * dsr;
    {
        const double ex[3] = {wDj[0] + 2.0 * Wv[0] + hd * vx, wDj[1] + 2.0 * Wv[1] + hd * vy,
                              wDj[2] + 2.0 * Wv[2] + hd * vz};
        double y[6];
        rot_vec(ex, fr, y);
#pragma unroll
        for (int k = 0; k < 3; ++k) rw[(R_EX + k) * ld] = y[k];
        rot_sym(wQj, fr, y);
#pragma unroll
        for (int k = 0; k < 6; ++k) rw[(R_WQ + k) * ld] = y[k];
    }
    rw[R_HSW * ld] = (j == i ? 0.5 : 1.0) * 0.5 * hs * wH;
}



// Primitive pairs [p0, p1) of one task of type (LI, LJ): accumulate the
// contracted derivative a3 = sum cc w.dM/dR_i, the overlap s and the i-centred
// dipole moments dacc.
template <int LI, int LJ, bool EX>
__device__ __forceinline__ void chunk_run(const double *__restrict__ pp, const double *rec, int ld,
                                          const double vec[3], int ki, int kj, int p0, int p1,
                                          double a3[3], double &s, double dacc[3])
{
    const int ei[3] = {cart_exp(LI, ki, 0), cart_exp(LI, ki, 1), cart_exp(LI, ki, 2)};
    const int ej[3] = {cart_exp(LJ, kj, 0), cart_exp(LJ, kj, 1), cart_exp(LJ, kj, 2)};
    double w[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) w[k] = rec[(R_W + k) * ld];
    const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
    // entries are loaded one iteration ahead
    double nx[TBX_PP_N];
#pragma unroll
    for (int f = 0; f < TBX_PP_N; ++f) nx[f] = pp[TBX_PP_N * p0 + f];
    for (int p = p0; p < p1; ++p) {
        const double mu = nx[TBX_PP_MU], ajo = nx[TBX_PP_AJO], aio = nx[TBX_PP_AIO];
        const double h = nx[TBX_PP_H], ai = nx[TBX_PP_AI], pre = nx[TBX_PP_PRE];
        if (p + 1 < p1) {
#pragma unroll
            for (int f = 0; f < TBX_PP_N; ++f) nx[f] = pp[TBX_PP_N * (p + 1) + f];
        }
        const double cc = exp(-(mu * r2)) * pre;
        const Mom M = make_mom(h);
        double V[3][3], G[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            grad1d_t<LI, LJ>(ei[k], ej[k], -vec[k] * ajo, vec[k] * aio, ai, M, V[k], G[k]);
        if (EX) {
            s += cc * (V[0][0] * V[1][0] * V[2][0]);
            dacc[0] += cc * (V[0][1] * V[1][0] * V[2][0]);
            dacc[1] += cc * (V[0][0] * V[1][1] * V[2][0]);
            dacc[2] += cc * (V[0][0] * V[1][0] * V[2][1]);
        }
        double c[3];
        contract_grad(G, V, w, c);
        a3[0] += cc * c[0];
        a3[1] += cc * c[1];
        a3[2] += cc * c[2];
    }
}

// The explicit v terms of one Cartesian pair and primitive pair (weight cc)
// in the s/p items, from its 1D factors V (moments 0..2 of each direction)
// and task record r: S ex + 2 Wj D_i into ex and cc S hsw added to ds.
__device__ __forceinline__ void explicit_terms(const double V[3][3], const double *r, int ld,
                                               double cc, double ex[3], double &ds)
{
    const double S = V[0][0] * V[1][0] * V[2][0];
    const double D0 = V[0][1] * V[1][0] * V[2][0];
    const double D1 = V[0][0] * V[1][1] * V[2][0];
    const double D2 = V[0][0] * V[1][0] * V[2][1];
    const double *wq = r + R_WQ * ld;
    const double wq0 = wq[0], wq1 = wq[ld], wq2 = wq[2 * ld], wq3 = wq[3 * ld],
                 wq4 = wq[4 * ld], wq5 = wq[5 * ld];
    const double WD0 = wq0 * D0 + 0.5 * (wq1 * D1 + wq3 * D2);
    const double WD1 = wq2 * D1 + 0.5 * (wq1 * D0 + wq4 * D2);
    const double WD2 = wq5 * D2 + 0.5 * (wq3 * D0 + wq4 * D1);
    ex[0] = S * r[R_EX * ld] + 2.0 * WD0;
    ex[1] = S * r[(R_EX + 1) * ld] + 2.0 * WD1;
    ex[2] = S * r[(R_EX + 2) * ld] + 2.0 * WD2;
    ds += cc * S * r[R_HSW * ld];
}

// 1D overlap table of one direction for a primitive pair (normalised
// moments): O[a][b] = int (y+pa)^a (y+pb)^b for a < NA, b < NB, from the
// recurrence O(a+1,0) = pa O(a,0) + a h O(a-1,0) (h = 1/(2 eta)) and the
// transfer O(a,b+1) = O(a+1,b) + (pb-pa) O(a,b).
template <int NA, int NB>
__device__ __forceinline__ void os_table(double pa, double pb, double h,
                                         double O[NA + NB - 1][NB])
{
    constexpr int N0 = NA + NB - 1;
    O[0][0] = 1.0;
    if (N0 > 1) O[1][0] = pa;
#pragma unroll
    for (int a = 1; a + 1 < N0; ++a) O[a + 1][0] = pa * O[a][0] + (a * h) * O[a - 1][0];
    const double d = pb - pa;
#pragma unroll
    for (int b = 0; b + 1 < NB; ++b)
#pragma unroll
        for (int a = 0; a + 1 < N0 - b; ++a) O[a][b + 1] = O[a + 1][b] + d * O[a][b];
}

// Primitive pairs [p0, p1) of an s/p shell pair (LI, LJ <= 1): the 1D tables
// are built once per primitive pair and shared by the Cartesian pairs of one
// row ki = ROW (ROW < 0: every row), with compile-time exponents; each
// Cartesian pair reads its task record (rec + cp, rec = the row's first
// Cartesian pair) and adds its contracted derivative to g, with EX the
// explicit v terms too, and its H0 share to ds.
template <int LI, int LJ, int ROW, bool EX>
__device__ __forceinline__ void shellpair_run(const double *__restrict__ pp, const double *rec,
                                              int ld, const double vec[3], int p0, int p1,
                                              double g[3], double &ds)
{
    constexpr int NCI = (LI + 1) * (LI + 2) / 2, NCJ = (LJ + 1) * (LJ + 2) / 2;
    constexpr int NA = LI + 4, NB = LJ + 1;
    const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
    for (int p = p0; p < p1; ++p) {
        const double *e = pp + TBX_PP_N * p;
        const double mu = e[TBX_PP_MU], ajo = e[TBX_PP_AJO], aio = e[TBX_PP_AIO];
        const double h = e[TBX_PP_H], ai = e[TBX_PP_AI], pre = e[TBX_PP_PRE];
        const double cc = exp(-(mu * r2)) * pre;
        const double ta = 2.0 * ai;
        double O[3][NA + NB - 1][NB];
#pragma unroll
        for (int k = 0; k < 3; ++k) os_table<NA, NB>(-vec[k] * ajo, vec[k] * aio, h, O[k]);
#pragma unroll
        for (int ki = (ROW < 0 ? 0 : ROW); ki < (ROW < 0 ? NCI : ROW + 1); ++ki)
#pragma unroll
            for (int kj = 0; kj < NCJ; ++kj) {
                // one Cartesian pair at a time: keeps the scheduler from
                // hoisting every pair's record loads (register pressure)
                __builtin_amdgcn_sched_barrier(0);
                const double *r = rec + ((ROW < 0 ? ki : 0) * NCJ + kj);
                double V[3][3], G[3][3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int a = cart_exp(LI, ki, k), b = cart_exp(LJ, kj, k);
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        V[k][m] = O[k][a + m][b];
                        G[k][m] = (a + m > 0) ? ta * O[k][a + m + 1][b] -
                                                    (double)(a + m) * O[k][a + m - 1][b]
                                              : ta * O[k][a + m + 1][b];
                    }
                }
                double w[10];
#pragma unroll
                for (int q = 0; q < 10; ++q) w[q] = r[(R_W + q) * ld];
                double cg[3];
                contract_grad(G, V, w, cg);
                const double c0 = cg[0], c1 = cg[1], c2 = cg[2];
                if (EX) {
                    double ex[3];
                    explicit_terms(V, r, ld, cc, ex, ds);
                    g[0] += cc * (c0 + ex[0]);
                    g[1] += cc * (c1 + ex[1]);
                    g[2] += cc * (c2 + ex[2]);
                } else {
                    g[0] += cc * c0;
                    g[1] += cc * c1;
                    g[2] += cc * c2;
                }
            }
    }
}

// Primitive pairs [p0, p1) of one row ki = row of a p-p shell pair, in the
// frame of the row (axis s along (row + s) mod 3, rot_vec): there the bra's
// exponent is 1 on axis 0 and 0 on the others, and the ket's is 1 on axis t
// for the Cartesian pair in slot t of the row's records (task_setup writes
// them in this frame), so every exponent is a compile-time constant.  The
// derivatives are rotated back into g at the end.
template <bool EX>
__device__ __forceinline__ void pprow_run(const double *__restrict__ pp, const double *rec,
                                          int ld, const double vec[3], int row, int p0, int p1,
                                          double g[3], double &ds)
{
    constexpr int NA = 5, NB = 2;
    double pv[3];
    rot_vec(vec, row, pv);
    const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
    double gs[3] = {0.0, 0.0, 0.0};
    for (int p = p0; p < p1; ++p) {
        const double *e = pp + TBX_PP_N * p;
        const double mu = e[TBX_PP_MU], ajo = e[TBX_PP_AJO], aio = e[TBX_PP_AIO];
        const double h = e[TBX_PP_H], ai = e[TBX_PP_AI], pre = e[TBX_PP_PRE];
        const double cc = exp(-(mu * r2)) * pre;
        const double ta = 2.0 * ai;
        // per axis: moments m = 0..2 and derivatives for the ket exponent b
        double V0[3][2][3], G0[3][2][3];  // [axis][b][m]
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double O[NA + NB - 1][NB];
            os_table<NA, NB>(-pv[k] * ajo, pv[k] * aio, h, O);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    if (k == 0) {  // bra exponent 1
                        V0[k][b][m] = O[m + 1][b];
                        G0[k][b][m] = ta * O[m + 2][b] - (double)(m + 1) * O[m][b];
                    } else {       // bra exponent 0
                        V0[k][b][m] = O[m][b];
                        G0[k][b][m] = (m > 0) ? ta * O[m + 1][b] - (double)m * O[m - 1][b]
                                              : ta * O[1][b];
                    }
                }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            __builtin_amdgcn_sched_barrier(0);
            const double *r = rec + t;
            double V[3][3], G[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    V[k][m] = V0[k][k == t ? 1 : 0][m];
                    G[k][m] = G0[k][k == t ? 1 : 0][m];
                }
            double w[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) w[q] = r[(R_W + q) * ld];
            double cg[3];
            contract_grad(G, V, w, cg);
            if (EX) {
                double ex[3];
                explicit_terms(V, r, ld, cc, ex, ds);
                gs[0] += cc * (cg[0] + ex[0]);
                gs[1] += cc * (cg[1] + ex[1]);
                gs[2] += cc * (cg[2] + ex[2]);
            } else {
                gs[0] += cc * cg[0];
                gs[1] += cc * cg[1];
                gs[2] += cc * cg[2];
            }
        }
    }
    // axis s of the frame is (row + s) mod 3
    g[0] += rot3(gs[0], gs[2], gs[1], row);
    g[1] += rot3(gs[1], gs[0], gs[2], row);
    g[2] += rot3(gs[2], gs[1], gs[0], row);
}

// One step of a segmented inclusive scan of (g, ds) over lanes with equal
// key: add the values of the lane that the DPP control CTRL (row mask
// ROWMASK) reads from, if its key is the same.  Lanes the move does not reach
// read INT_MIN, which no key equals.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_scan_step(int key, double g[3], double &ds)
{
    const int k2 = __builtin_amdgcn_update_dpp(INT_MIN, key, CTRL, ROWMASK, 0xf, false);
    const double y0 = dpp_move<CTRL, ROWMASK>(g[0]), y1 = dpp_move<CTRL, ROWMASK>(g[1]);
    const double y2 = dpp_move<CTRL, ROWMASK>(g[2]), y3 = dpp_move<CTRL, ROWMASK>(ds);
    if (k2 == key) {
        g[0] += y0;
        g[1] += y1;
        g[2] += y2;
        ds += y3;
    }
}

// Segmented inclusive scan of N doubles over the 64 lanes of a wavefront,
// runs of equal key (the same DPP steps as the work-item sums): the last
// lane of each run ends with the run's total.
template <int CTRL, int ROWMASK, int N>
__device__ __forceinline__ void seg_scan_step_n(int key, double (&v)[N])
{
    const int k2 = __builtin_amdgcn_update_dpp(INT_MIN, key, CTRL, ROWMASK, 0xf, false);
    double y[N];
#pragma unroll
    for (int n = 0; n < N; ++n) y[n] = dpp_move<CTRL, ROWMASK>(v[n]);
    if (k2 == key) {
#pragma unroll
        for (int n = 0; n < N; ++n) v[n] += y[n];
    }
}

template <int N>
__device__ __forceinline__ void seg_scan_n(int key, double (&v)[N])
{
    seg_scan_step_n<0x111, 0xf>(key, v);  // row_shr:1
    seg_scan_step_n<0x112, 0xf>(key, v);  // row_shr:2
    seg_scan_step_n<0x114, 0xf>(key, v);  // row_shr:4
    seg_scan_step_n<0x118, 0xf>(key, v);  // row_shr:8
    seg_scan_step_n<0x142, 0xa>(key, v);  // row_bcast:15 into rows 1 and 3
    seg_scan_step_n<0x143, 0xc>(key, v);  // row_bcast:31 into rows 2 and 3
}

// sum over the 64 lanes of a wavefront (every lane gets the total)
__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// One workgroup per work unit (up to TBX_UNIT_PAIRS atom pairs of one
// molecule, at most TBX_UNIT_TASKS Cartesian-pair tasks).  (1) The unit's
// pairs are staged in LDS and their shell pairs laid out by type (l_i, l_j);
// (2) one lane per task builds the task record and counts its kept primitive
// pairs; (3) the kept pairs are cut into chunks of A.chunk and a block scan
// turns them into a dense list of work items; (4) every lane runs work items
// (type-major, so a wavefront mostly runs one specialised code path) and adds
// the results to its pair's LDS sums; (5) the pair sums go to atoms j, the
// segment sums to atoms i and the unit sum to sigma, with FP64 atomics.
template <int MODE, bool EX>
__device__ __forceinline__ void gradient_unit(const GradArgs &A, int unit)
{
    // without EX a task record holds the 10 moment weights only, and the
    // shell-pair factors live in the records' space until the records are
    // written (3 * GTHREADS <= 10 * GTHREADS doubles)
    constexpr int nrec = EX ? NREC : R_EX;
    __shared__ GradStage U;
    __shared__ double rec[nrec * GTHREADS];
    __shared__ double accP[TBX_GRAD_UNIT_PAIRS][3 + 2 * MAXSH];  // G (3), dse_i, dse_j per shell
    __shared__ int tinfo[GTHREADS];  // type | ki << 4 | kj << 8 | pair << 12 | ish << 24 | jsh << 27
    __shared__ int tpp[GTHREADS];    // first primitive pair of the task
    __shared__ int tkept[GTHREADS];
    __shared__ __align__(16) int coff[GTHREADS + 1];  // first chunk of each task
    __shared__ int spk[GTHREADS], spp[GTHREADS];  // kept pairs, first entry
    __shared__ double sph_own[EX ? 3 : 1][GTHREADS];  // hs, sesum, dsr per shell pair
    double (*sph)[GTHREADS] = EX ? sph_own : reinterpret_cast<double (*)[GTHREADS]>(rec);
    const TbxSpecies &sp = A.sp;
    const TbxMols &mt = A.mol;
    const int tid = threadIdx.x;
    const TbxUnitHead H = mt.unit[unit];
    const int npair = H.counts & 255, nseg = (H.counts >> 8) & 255, nspair = H.counts >> 16;
    const int m = H.mol;
    const int nao = H.nao;
    const long long dense0 = H.dense0;
    const int ao_m = H.ao_m;
    const int chunk = A.chunk;
    for (int k = tid; k < TBX_GRAD_UNIT_PAIRS * (3 + 2 * MAXSH); k += GTHREADS) (&accP[0][0])[k] = 0.0;
    // (1) the unit's plan
    load_unit(H, mt.plan, U);
    __syncthreads();

    // (2) pair vectors.  A unit with no pair within the cutoff has no work
    // (the host groups each molecule's distant pairs into units of their own
    // when given a geometry hint): it ends here, before any output.
    bool live = false;
    if (tid < npair) {
        double v[3];
        pair_vec(mt, A.xyz, U.pinf[tid], U.pext[tid], v);
#pragma unroll
        for (int c = 0; c < 3; ++c) U.pv[tid][c] = v[c];
        const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        live = r2 >= kEps && r2 <= sp.cutoff2;
    }
    if (!__syncthreads_or(live)) return;
    // shell pairs: screening counts and H0 factors (one lane per shell pair)
    for (int k = tid; k < nspair; k += GTHREADS) {
        int kept, pi0;
        double hs, sesum, dsr;
        shellpair_setup(sp, mt, A.xyz, U, A.se, k, false, kept, pi0, hs, sesum, dsr);
        spk[k] = kept;
        spp[k] = pi0;
        sph[0][k] = hs;
        sph[1][k] = sesum;
        sph[2][k] = dsr;
    }
    __syncthreads();
    // task records: one Cartesian pair per lane (ntask <= GTHREADS)
    const int ntask = U.toff[9];
    const int lane = tid & 63;
    int nch = 0;
    int spx = 0, kept = 0;
    int4 d = make_int4(0, 0, 0, 0), e = make_int4(0, 0, 0, 0);
    double hs = 0.0, sesum = 0.0, dsr = 0.0;
    if (tid < ntask) {
        d = decode_task(U, tid, spx);
        e = unpack_spl(U.spl[spx]);
        kept = spk[spx];
        hs = sph[0][spx];
        sesum = sph[1][spx];
        dsr = sph[2][spx];
    }
    if (!EX) __syncthreads();  // the shell-pair factors are read before the records overwrite them
    int xkey = -1 - lane;  // shell pair of the lane's explicit terms; unique if idle
    double gx[3] = {0.0, 0.0, 0.0}, dsx = 0.0;
    if (tid < ntask) {
        if (kept > 0) {
            const PairInfo pr = U.pinf[e.x];
            const double v[3] = {U.pv[e.x][0], U.pv[e.x][1], U.pv[e.x][2]};
            const int li = d.x / 3, lj = d.x % 3;
            task_setup<MODE, EX>(A, pr, v, e.z, e.y, li, lj, d.y, d.z, nao, dense0, ao_m, hs,
                                 sesum, dsr, rec + tid, GTHREADS, gx, dsx);
            // with a d shell, the spherical element (ki, kj) of the shell pair
            // is this lane's
            if (!EX && (li == 2 || lj == 2) && d.y < 2 * li + 1 && d.z < 2 * lj + 1) {
                const int mu = pr.aoj + ao_first(pr.dj, e.y) + d.z;
                const int nu = pr.aoi + ao_first(pr.di, e.z) + d.y;
                explicit_element<MODE>(A, dense0 + (long long)nu * nao + mu,
                                       dense0 + (long long)mu * nao + nu, pr.i, pr.j, v, hs,
                                       sesum, dsr, gx, dsx);
            }
            xkey = e.x | (e.z << 12) | (e.y << 15);
        }
        tinfo[tid] = d.x | (d.y << 4) | (d.z << 8) | (e.x << 12) | (e.z << 24) | (e.y << 27);
        tpp[tid] = spp[spx];
        tkept[tid] = kept;
        // s/p shell pairs run whole on their first Cartesian pair (p-p: by row,
        // on the first Cartesian pair of each row)
        const bool whole = (d.x == 0 || d.x == 1 || d.x == 3);
        const bool rows = (d.x == 4);
        nch = ((whole && d.w != 0) || (rows && d.z != 0)) ? 0 : (kept + chunk - 1) / chunk;
    }
    if (!EX) {
        // the explicit terms of a shell pair's elements (consecutive lanes)
        // summed across the wavefront, then added to the pair's sums
        seg_scan_step<0x111, 0xf>(xkey, gx, dsx);
        seg_scan_step<0x112, 0xf>(xkey, gx, dsx);
        seg_scan_step<0x114, 0xf>(xkey, gx, dsx);
        seg_scan_step<0x118, 0xf>(xkey, gx, dsx);
        seg_scan_step<0x142, 0xa>(xkey, gx, dsx);
        seg_scan_step<0x143, 0xc>(xkey, gx, dsx);
        const int knext = __shfl_down(xkey, 1, 64);
        if (xkey >= 0 && (lane == 63 || knext != xkey)) {
            const int p = xkey & 4095, ish = (xkey >> 12) & 7, jsh = (xkey >> 15) & 7;
            atomicAdd(&accP[p][0], gx[0]);
            atomicAdd(&accP[p][1], gx[1]);
            atomicAdd(&accP[p][2], gx[2]);
            atomicAdd(&accP[p][3 + ish], dsx);
            atomicAdd(&accP[p][3 + MAXSH + jsh], dsx);
        }
    }
    // (3) chunk offsets: block-wide exclusive scan
    const int nchunk = block_scan<GTHREADS>(nch, coff, U.wsum);

    // (4) work items, one per lane and step; the results of a shell pair's
    // items (consecutive lanes) are summed across the wavefront before the
    // LDS atomics
    for (int base = 0; base < nchunk; base += GTHREADS) {
        const int c = base + tid;
        int key = -1 - lane;  // shell pair of the item (p | ish << 12 | jsh << 15); unique if idle
        double g[3] = {0, 0, 0}, ds = 0.0;
        if (c < nchunk) {
        const int tk = find_task<GTHREADS>(coff, c);
        const int info = tinfo[tk];
        const int t = info & 15, ki = (info >> 4) & 15, kj = (info >> 8) & 15;
        const int p = (info >> 12) & 4095;
        key = info >> 12;
        const int q = c - coff[tk];
        const int p0 = tpp[tk] + q * chunk;
        const int p1 = tpp[tk] + min(tkept[tk], (q + 1) * chunk);
        const double *r = rec + tk;
        const double vec[3] = {U.pv[p][0], U.pv[p][1], U.pv[p][2]};
        const int L = GTHREADS;
        if (t == 0 || t == 1 || t == 3 || t == 4) {
            switch (t) {
            case 0: shellpair_run<0, 0, -1, EX>(sp.pp, r, L, vec, p0, p1, g, ds); break;
            case 1: shellpair_run<0, 1, -1, EX>(sp.pp, r, L, vec, p0, p1, g, ds); break;
            case 3: shellpair_run<1, 0, -1, EX>(sp.pp, r, L, vec, p0, p1, g, ds); break;
            default: pprow_run<EX>(sp.pp, r, L, vec, ki, p0, p1, g, ds); break;  // p-p by rows
            }
        } else {
            double a3[3] = {0, 0, 0}, dacc[3] = {0, 0, 0}, s = 0.0;
            switch (t) {
            case 2: chunk_run<0, 2, EX>(sp.pp, r, L, vec, ki, kj, p0, p1, a3, s, dacc); break;
            case 5: chunk_run<1, 2, EX>(sp.pp, r, L, vec, ki, kj, p0, p1, a3, s, dacc); break;
            case 6: chunk_run<2, 0, EX>(sp.pp, r, L, vec, ki, kj, p0, p1, a3, s, dacc); break;
            case 7: chunk_run<2, 1, EX>(sp.pp, r, L, vec, ki, kj, p0, p1, a3, s, dacc); break;
            default: chunk_run<2, 2, EX>(sp.pp, r, L, vec, ki, kj, p0, p1, a3, s, dacc); break;
            }
            if (!EX) {
                g[0] = a3[0];
                g[1] = a3[1];
                g[2] = a3[2];
            } else {
            // explicit v dependence: S ex + 2 Wj D_i
            const double *wq = r + R_WQ * L;
            const double wq0 = wq[0], wq1 = wq[L], wq2 = wq[2 * L], wq3 = wq[3 * L],
                         wq4 = wq[4 * L], wq5 = wq[5 * L];
            const double WD0 = wq0 * dacc[0] + 0.5 * (wq1 * dacc[1] + wq3 * dacc[2]);
            const double WD1 = wq2 * dacc[1] + 0.5 * (wq1 * dacc[0] + wq4 * dacc[2]);
            const double WD2 = wq5 * dacc[2] + 0.5 * (wq3 * dacc[0] + wq4 * dacc[1]);
            g[0] = a3[0] + s * r[R_EX * L] + 2.0 * WD0;
            g[1] = a3[1] + s * r[(R_EX + 1) * L] + 2.0 * WD1;
            g[2] = a3[2] + s * r[(R_EX + 2) * L] + 2.0 * WD2;
            ds = s * r[R_HSW * L];
            }
        }
        }
        // segmented inclusive scan over runs of equal key (a shell pair's
        // items are consecutive lanes) with DPP moves: shifts by 1, 2, 4, 8
        // within rows of 16 lanes, then the row broadcasts of lanes 15 and 31;
        // the last lane of each run ends with the run's total
        // (without EX the items have no dE/dse share)
        if (EX) {
            seg_scan_step<0x111, 0xf>(key, g, ds);  // row_shr:1
            seg_scan_step<0x112, 0xf>(key, g, ds);  // row_shr:2
            seg_scan_step<0x114, 0xf>(key, g, ds);  // row_shr:4
            seg_scan_step<0x118, 0xf>(key, g, ds);  // row_shr:8
            seg_scan_step<0x142, 0xa>(key, g, ds);  // row_bcast:15 into rows 1 and 3
            seg_scan_step<0x143, 0xc>(key, g, ds);  // row_bcast:31 into rows 2 and 3
        } else {
            seg_scan_n<3>(key, g);
        }
        const int knext = __shfl_down(key, 1, 64);
        if (key >= 0 && (lane == 63 || knext != key)) {
            const int p = key & 4095, ish = (key >> 12) & 7, jsh = (key >> 15) & 7;
            atomicAdd(&accP[p][0], g[0]);
            atomicAdd(&accP[p][1], g[1]);
            atomicAdd(&accP[p][2], g[2]);
            if (EX) {
                atomicAdd(&accP[p][3 + ish], ds);
                atomicAdd(&accP[p][3 + MAXSH + jsh], ds);
            }
        }
    }
    __syncthreads();

    // (5) one lane per pair (npair <= TBX_GRAD_UNIT_PAIRS: the first two
    // wavefronts).  The pair's sums go to atom j (HBM atomics).  The atom-i
    // sums of each segment (a run of consecutive pairs with one row atom i)
    // are reduced by a segmented scan across the wavefront and added by the
    // segment's last lane in the wavefront (a segment across the wavefront
    // boundary adds its two parts); each wavefront's share of the unit's
    // sigma is a wavefront sum added by nine lanes.
    if (tid < TBX_GRAD_UNIT_PAIRS) {
        const bool on = tid < npair;
        const int key = on ? (U.pext[tid] & 31) : 64 + tid;  // segment; unique on idle lanes
        double gi[3 + MAXSH], sg[6];
#pragma unroll
        for (int q = 0; q < 3 + MAXSH; ++q) gi[q] = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) sg[q] = 0.0;
        PairInfo pr{};
        if (on) {
            pr = U.pinf[tid];
            const bool self = (pr.i == pr.j);
            const double G0 = accP[tid][0], G1 = accP[tid][1], G2 = accP[tid][2];
            const double vx = U.pv[tid][0], vy = U.pv[tid][1], vz = U.pv[tid][2];
            if (!self && (G0 != 0.0 || G1 != 0.0 || G2 != 0.0)) {
                unsafeAtomicAdd(&A.grad[3 * pr.j + 0], -G0);
                unsafeAtomicAdd(&A.grad[3 * pr.j + 1], -G1);
                unsafeAtomicAdd(&A.grad[3 * pr.j + 2], -G2);
                gi[0] = G0;
                gi[1] = G1;
                gi[2] = G2;
            }
            const double fs = self ? 0.25 : 0.5;
            sg[0] = fs * 2.0 * vx * G0;
            sg[1] = fs * (vx * G1 + G0 * vy);
            sg[2] = fs * 2.0 * vy * G1;
            sg[3] = fs * (vx * G2 + G0 * vz);
            sg[4] = fs * (vy * G2 + G1 * vz);
            sg[5] = fs * 2.0 * vz * G2;
            for (int q = 0; q < (pr.dj & 7); ++q)
                if (accP[tid][3 + MAXSH + q] != 0.0)
                    unsafeAtomicAdd(&A.dse[pr.shj + q], accP[tid][3 + MAXSH + q]);
#pragma unroll
            for (int q = 0; q < MAXSH; ++q) gi[3 + q] = accP[tid][3 + q];
        }
        seg_scan_n<3 + MAXSH>(key, gi);
#pragma unroll
        for (int q = 0; q < 6; ++q) sg[q] = wave_sum_d(sg[q]);
        const int knext = __shfl_down(key, 1, 64);
        if (on && (tid == npair - 1 || knext != key || (tid & 63) == 63)) {  // segment's last lane
            for (int q = 0; q < 3; ++q)
                if (gi[q] != 0.0) unsafeAtomicAdd(&A.grad[3 * pr.i + q], gi[q]);
            for (int q = 0; q < (pr.di & 7); ++q)
                if (gi[3 + q] != 0.0) unsafeAtomicAdd(&A.dse[pr.shi + q], gi[3 + q]);
        }
        if ((tid & 63) < 9) {
            double v = 0.0;
#pragma unroll
            for (int k9 = 0; k9 < 9; ++k9) v = ((tid & 63) == k9) ? sg[kSigPack[k9]] : v;
            if (v != 0.0) unsafeAtomicAdd(&A.sigma[9 * m + (tid & 63)], v);
        }
    }
}

// EX: the explicit pair-vector terms in the work items (the forward's S and
// D not given, or lattice images).  Without them the kernel fits three
// wavefronts per SIMD: 168 VGPRs and under 53 KB of LDS per 256-thread
// workgroup (3 workgroups per CU).
template <int MODE, bool EX>
__global__ void __launch_bounds__(GTHREADS, EX ? 2 : 3) k_gradient(GradArgs A)
{
    gradient_unit<MODE, EX>(A, xcd_unit(blockIdx.x, gridDim.x));  // XCD-contiguous
}

// The far gradient units (no pair within cutoff + skin at the pair-order
// hint): nothing to do while every atom is within skin/2 of its hint
// position (flag 0, set by k_hint_check), since then every such pair is
// still beyond the cutoff.  Otherwise the workgroups walk them.
template <int MODE, bool EX>
__global__ void __launch_bounds__(GTHREADS, EX ? 2 : 3) k_gradient_far(GradArgs A,
                                                                     const int *stale)
{
    if (*stale == 0) return;
    for (int u = A.mol.nunit_near + blockIdx.x; u < A.mol.nunit; u += gridDim.x) {
        __syncthreads();  // the previous unit's LDS is read until its last phase
        gradient_unit<MODE, EX>(A, u);
    }
}

// stale = 1 if any atom moved more than sqrt(hint_r2) from its hint position
__global__ void k_hint_check(TbxMols mt, const double *xyz, int *stale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (i < mt.nat) {
        const double dx = xyz[3 * i] - mt.hint[3 * i], dy = xyz[3 * i + 1] - mt.hint[3 * i + 1],
                     dz = xyz[3 * i + 2] - mt.hint[3 * i + 2];
        moved = !(dx * dx + dy * dy + dz * dz <= mt.hint_r2);  // NaN counts as moved
    }
    if (__any(moved) && (threadIdx.x & 63) == 0) atomicOr(stale, 1);
}

__global__ void k_fill_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// On-site blocks of the adjoint: H = S (se_i + se_j)/2 contributes to dE/dse
// only.  16 lanes per atom, looping over its Cartesian pairs.
__global__ void __launch_bounds__(NTHREADS) k_onsite(GradArgs A)
{
    const TbxSpecies &sp = A.sp;
    const TbxMols &mt = A.mol;
    const int i = blockIdx.x * (NTHREADS / 16) + (threadIdx.x >> 4);
    if (i >= mt.nat) return;
    const int spi = mt.at_sp[i];
    const int nci = sp.ncart[spi];
    const int m = mt.at_mol[i];
    const int nao = mt.mol_nao[m];
    const long long dense0 = mt.mol_dense[m];
    for (int tk = threadIdx.x & 15; tk < nci * nci; tk += 16) {
        const int ci = tk / nci, cj = tk - ci * nci;
        const int ei_ = sp.cart_exp[spi * TBX_MAXCART + ci];
        const int ej_ = sp.cart_exp[spi * TBX_MAXCART + cj];
        const int ish = sp.cart_sh[spi * TBX_MAXCART + ci];
        const int jsh = sp.cart_sh[spi * TBX_MAXCART + cj];
        const int ki = ei_ >> 12, kj = ej_ >> 12;
        const int sloti = ish + sp.msh * spi, slotj = jsh + sp.msh * spi;
        const int li = sp.ang[sloti], lj = sp.ang[slotj];
        const int muj0 = mt.at_ao[i] + sp.sh_ao[slotj];
        const int nui0 = mt.at_ao[i] + sp.sh_ao[sloti];
        double wHon = 0.0;
        for (int a = 0; a < cart_n(lj, kj); ++a)
            for (int b = 0; b < cart_n(li, ki); ++b)
                wHon += cart_w(lj, kj, a) * cart_w(li, ki, b) *
                        A.Hb[dense0 + (long long)(nui0 + cart_m(li, ki, b)) * nao + muj0 +
                             cart_m(lj, kj, a)];
        if (wHon == 0.0) continue;
        // the same-centre overlap of the pair (a species constant)
        const double s = sp.onsite[((spi * TBX_MAXCART + ci) * TBX_MAXCART + cj) * 10];
        if (s != 0.0) {
            unsafeAtomicAdd(&A.dse[mt.at_sh[i] + ish], 0.5 * wHon * s);
            unsafeAtomicAdd(&A.dse[mt.at_sh[i] + jsh], 0.5 * wHon * s);
        }
    }
}

// Neighbour-list check of the single-structure entry points: pairs (j <= i,
// image) within the cutoff must match the caller's half-list counts.
__global__ void k_listcheck(TbxSpecies sp, TbxMols mt, const double *xyz, const int *nnl,
                            int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mt.nat) return;
    const int m = mt.at_mol[i];
    const int a0 = mt.mol_atom[m];
    const int2 tr = mt.mol_tr[m];
    int n = 0;
    for (int j = a0; j <= i; ++j)
        for (int itr = tr.x; itr < tr.x + tr.y; ++itr) {
            const double vx = xyz[3 * i] - xyz[3 * j] - mt.trans[3 * itr];
            const double vy = xyz[3 * i + 1] - xyz[3 * j + 1] - mt.trans[3 * itr + 1];
            const double vz = xyz[3 * i + 2] - xyz[3 * j + 2] - mt.trans[3 * itr + 2];
            const double r2 = vx * vx + vy * vy + vz * vz;
            if (r2 >= kEps && r2 <= sp.cutoff2) ++n;
        }
    if (n != nnl[i]) atomicOr(flag, 1);
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

extern "C" hipError_t tbx_launch_fill(double *p, long long n, double v, hipStream_t stream);

extern "C" hipError_t tbx_launch_forward(const TbxSpecies *sp, const TbxMols *mol,
                                         const double *xyz, const double *se, double *S, double *D,
                                         double *Q, double *H, long long ndense, const int *nnl,
                                         int *flag, int chunk, hipStream_t stream)
{
    const int accumulate = mol->ntrans > 1;
    FwdArgs a{*sp, *mol, xyz, se, S, D, Q, H, accumulate, chunk > 0 ? chunk : FWD_CHUNK_DEFAULT,
              flag};
    (void)hipGetLastError();  // launch errors are read below; drop stale ones
    if (mol->nfunit_far > 0) {  // does the pair-order hint still hold for the far units?
        if (nnl) return hipErrorInvalidValue;  // the flag serves one purpose per launch
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, flag, 0);
        hipLaunchKernelGGL(k_hint_check, dim3((mol->nat + 255) / 256), dim3(256), 0, stream,
                           *mol, xyz, flag);
    }
    if (accumulate) {  // lattice images add into the same elements
        double *outs[4] = {S, H, D, Q};
        const long long n[4] = {ndense, ndense, 3 * ndense, 6 * ndense};
        for (int k = 0; k < 4; ++k) {
            const hipError_t e = tbx_launch_fill(outs[k], n[k], 0.0, stream);
            if (e != hipSuccess) return e;
        }
    }
    if (mol->nfunit > 0) {
        if (sp->kdiat)
            hipLaunchKernelGGL(k_forward<true>, dim3(mol->nfunit), dim3(NTHREADS), 0, stream, a);
        else
            hipLaunchKernelGGL(k_forward<false>, dim3(mol->nfunit), dim3(NTHREADS), 0, stream, a);
    }
    if (nnl)
        hipLaunchKernelGGL(k_listcheck, dim3((mol->nat + 255) / 256), dim3(256), 0, stream, *sp,
                           *mol, xyz, nnl, flag);
    return hipGetLastError();
}

extern "C" hipError_t tbx_launch_gradient(int mode, const TbxSpecies *sp, const TbxMols *mol,
                                          const double *xyz, const double *se,
                                          const double *P, const double *W, const double *vao,
                                          const double *vdp, const double *vqp, int nspin,
                                          long long pstride, long long vstride,
                                          const double *Sb, const double *Db, const double *Qb,
                                          const double *Hb, const double *Sf, const double *Df,
                                          double *grad, double *dse, double *sigma,
                                          const int *nnl, int *flag, int chunk,
                                          hipStream_t stream)
{
    // the per-element explicit terms need the forward's outputs and one
    // translation (T = 0) per structure: each element then belongs to one pair
    const bool ex = !(Sf && Df && mol->zero_image);
    GradArgs a{*sp, *mol, xyz, se, P, W, vao, vdp, vqp, nspin, pstride, vstride,
               Sb, Db, Qb, Hb, ex ? nullptr : Sf, ex ? nullptr : Df, grad, dse, sigma, nnl, flag,
               chunk > 0 ? chunk : GRAD_CHUNK_DEFAULT};
    (void)hipGetLastError();
    const int nfar = mol->nunit - mol->nunit_near;
    if (nfar > 0 && nnl) return hipErrorInvalidValue;  // the flag serves one purpose per launch
    if (nfar > 0) {  // is the pair-order hint still good for the far units?
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, flag, 0);
        hipLaunchKernelGGL(k_hint_check, dim3((mol->nat + 255) / 256), dim3(256), 0, stream,
                           *mol, xyz, flag);
    }
    const dim3 near(mol->nunit_near), far(nfar < 2048 ? nfar : 2048), blk(GTHREADS);
    const int *stale = flag;
    if (mol->nunit_near > 0) {
        if (mode == GRAD_PLAIN && ex)
            hipLaunchKernelGGL((k_gradient<GRAD_PLAIN, true>), near, blk, 0, stream, a);
        else if (mode == GRAD_PLAIN)
            hipLaunchKernelGGL((k_gradient<GRAD_PLAIN, false>), near, blk, 0, stream, a);
        else if (ex)
            hipLaunchKernelGGL((k_gradient<GRAD_ADJOINT, true>), near, blk, 0, stream, a);
        else
            hipLaunchKernelGGL((k_gradient<GRAD_ADJOINT, false>), near, blk, 0, stream, a);
    }
    if (nfar > 0) {
        if (mode == GRAD_PLAIN && ex)
            hipLaunchKernelGGL((k_gradient_far<GRAD_PLAIN, true>), far, blk, 0, stream, a, stale);
        else if (mode == GRAD_PLAIN)
            hipLaunchKernelGGL((k_gradient_far<GRAD_PLAIN, false>), far, blk, 0, stream, a, stale);
        else if (ex)
            hipLaunchKernelGGL((k_gradient_far<GRAD_ADJOINT, true>), far, blk, 0, stream, a, stale);
        else
            hipLaunchKernelGGL((k_gradient_far<GRAD_ADJOINT, false>), far, blk, 0, stream, a,
                               stale);
    }
    const int per = NTHREADS / 16;
    if (mode == GRAD_ADJOINT)
        hipLaunchKernelGGL(k_onsite, dim3((mol->nat + per - 1) / per), dim3(NTHREADS), 0, stream,
                           a);
    if (nnl)
        hipLaunchKernelGGL(k_listcheck, dim3((mol->nat + 255) / 256), dim3(256), 0, stream, *sp,
                           *mol, xyz, nnl, flag);
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

// Host self-test: the packed maps reproduce the tables (returns the number
// of mismatches; 0 on success).
extern "C" int tbx_selftest_maps(void)
{
    int bad = 0;
    for (int l = 0; l <= 2; ++l) {
        for (int m = 0; m < 2 * l + 1; ++m) {
            bad += sph_n(l, m) != kSphN[l][m];
            for (int a = 0; a < kSphN[l][m]; ++a)
                bad += (sph_c(l, m, a) != kSphC[l][m][a]) + (sph_w(l, m, a) != kSphW[l][m][a]);
        }
        for (int c = 0; c < (l + 1) * (l + 2) / 2; ++c) {
            bad += cart_n(l, c) != kCartN[l][c];
            for (int k = 0; k < kCartN[l][c]; ++k)
                bad += (cart_m(l, c, k) != kCartM[l][c][k]) + (cart_w(l, c, k) != kCartW[l][c][k]);
            for (int k = 0; k < 3; ++k) bad += cart_exp(l, c, k) != kCartExp[l][c][k];
        }
    }
    return bad;
}

// Streaming copy (traffic calibration of the hardware counters)
__global__ void k_copy(const double *__restrict__ src, double *__restrict__ dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        dst[k] = src[k];
}

extern "C" hipError_t tbx_launch_copy(const double *src, double *dst, long long n,
                                      hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, stream, src, dst, n);
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
