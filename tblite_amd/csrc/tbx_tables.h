// tbx_tables.h -- layout shared by the host set-up (tbx_capi.hip) and the
// gfx950 kernels (tbx_kernels.hip).  Internal, not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TBX_MAXCART 24 /* Cartesian functions per species: <= 4 shells, l <= 2 */
#define TBX_MAXAO 20   /* spherical functions per species */
#define TBX_MAXPRIM 2048 /* primitives over all shells of all species (staged in LDS) */

/* fields of a primitive-pair entry (ai, aj: exponents; ci, cj: contraction
   coefficients incl. normalisation) */
#define TBX_PP_MU 0   /* ai aj / (ai + aj) */
#define TBX_PP_AJO 1  /* aj / (ai + aj) */
#define TBX_PP_AIO 2  /* ai / (ai + aj) */
#define TBX_PP_H 3    /* 1 / (2 (ai + aj)) */
#define TBX_PP_AI 4   /* ai */
#define TBX_PP_PRE 5  /* pi^(3/2) (ai + aj)^(-3/2) ci cj */
#define TBX_PP_N 6

// Species tables (device pointers).  Shell slot = ish + msh*isp.
struct TbxSpecies {
    int nsp, msh;
    const int *nsh;       // [nsp]
    const int *ang;       // [nsp*msh]
    const int *nprim;     // [nsp*msh]
    const double *alpha;  // [nsp*msh*MAXG]
    const double *coeff;  // [nsp*msh*MAXG]
    const double *shpoly; // [nsp*msh]
    const double *rad;    // [nsp]
    const double *hscale; // [msh*msh*nsp*nsp], (jsh, ish, jsp, isp)
    const int *ncart;     // [nsp] Cartesian functions of the species
    const int *nao;       // [nsp] spherical functions of the species
    const int *cart_sh;   // [nsp*TBX_MAXCART] shell of each Cartesian function
    const int *cart_exp;  // [nsp*TBX_MAXCART] lx | ly<<4 | lz<<8 | (index in shell)<<12
    const int *ao_sh;     // [nsp*TBX_MAXAO]  shell of each spherical function
    const int *ao_m;      // [nsp*TBX_MAXAO]  m index (0..2l) within the shell
    const int *sh_cart;   // [nsp*msh] first Cartesian function of the shell within the atom
    const int *sh_ao;     // [nsp*msh] first spherical function of the shell within the atom
    const double *prim;   // [2*nprim_tot] (alpha, coeff) of every primitive, shells packed
    const int *prim_off;  // [nsp*msh] first primitive of the shell in prim
    int nprim_tot;
    // primitive pairs of every ordered shell-slot pair (si, sj), sorted by the
    // reduced exponent mu = ai aj / (ai + aj): entries pp_off[si*nslot+sj] ..
    // pp_off[si*nslot+sj+1], 6 doubles each (TBX_PP_* fields)
    int nslot;            // nsp*msh
    const int *pp_off;    // [nslot*nslot+1]
    const double *pp;     // [TBX_PP_N * entries]
    const double *pp_mu;  // [entries] reduced exponents alone (screening counts)
    // same-centre moments of the Cartesian pair (ci, cj) of a species, sums
    // over its primitive pairs (no geometry, no screening): S, D (x, y, z) and
    // Q (xx, xy, yy, xz, yz, zz) about the centre, the on-site blocks
    const double *onsite; // [nsp*TBX_MAXCART*TBX_MAXCART*10]
    const int *ao_info;   // [nsp*TBX_MAXAO] shell | m << 4 | l << 8 of each spherical function
    const int *ao_desc;   // [nsp] shells (bits 0-2) | l of shell s << (3 + 2 s)
    // diatomic-frame scaling of the overlap in H0: (ksig, kpi, kdel) of the
    // species pair (isp of atom i, jsp of atom j) at 3 * (isp + nsp * jsp),
    // or null (off)
    const double *kdiat;
    double intcut;
    double cutoff2;
};

// Work units: (i, j, image) atom pairs of one molecule, in segments j in
// [j0, j1) of one row atom i at lattice vector itr, packed so
// that a unit holds at most TBX_UNIT_TASKS Cartesian-pair tasks and
// TBX_UNIT_PAIRS atom pairs (forward units; TBX_GRAD_UNIT_* for the gradient)
// and TBX_UNIT_SEGS segments.  One workgroup runs one unit.
#define TBX_UNIT_TASKS 256       /* forward units (and workgroup size) */
#define TBX_UNIT_PAIRS 128
#define TBX_UNIT_SEGS 32
#define TBX_GRAD_UNIT_TASKS 256  /* gradient units (and workgroup size) */
#define TBX_GRAD_UNIT_PAIRS 128
struct TbxSeg {
    int i, j0, j1, itr;
};

// One atom pair of a unit
struct TbxPair {
    int i, j;      // atoms (global)
    int spi, spj;  // species
    int aoi, aoj;  // first AO of the atoms, molecule-local
    int shi, shj;  // first shell of the atoms (global)
    int di, dj;    // shell layouts of the species (TbxSpecies::ao_desc)
};

// Unit header (32 bytes).  The unit's plan -- everything about it that does not depend
// on the geometry -- is a run of 16-byte words at plan[16 * plan ...] in this
// order, each section padded to whole words:
//   pairs      npair TbxPair (40 bytes each)
//   pext       npair ints: lattice vector << 5 | segment
//   segments   nseg int4: first AO of the row atom i (the block's column), first AO
//              of the segment's atoms j (its rows), rows | mirrored rows << 16,
//              AOs of atom i (all molecule-local)
//   shell pairs nspair ints: pair | jsh << 8 | ish << 12 | type << 16, type-major
//              (type = 3 l_i + l_j)
//   offsets    toff[10] then tsp[10]: first task / shell pair of each type
// A forward unit writes, segment by segment, the segment's block A and its
// mirror (tbx_elem_locate): nel dense elements in all.
struct TbxUnitHead {
    long long dense0;  // first dense element of the unit's molecule
    int plan;          // first 16-byte word of the plan
    int counts;        // npair | nseg << 8 | nspair << 16
    int nao, ao_m;     // AOs of the molecule, its first AO (global)
    int mol;
    int nel;           // dense elements written (forward units) | TBX_UNIT_FAR
};
// TbxUnitHead::nel flag of a forward unit none of whose pairs lies within
// cutoff + skin at the pair-order hint; TBX_NEL_MASK gives the count
#define TBX_UNIT_FAR (1 << 30)
#define TBX_NEL_MASK 0xffff

// The dense elements of a forward unit, segment k after segment k - 1:
// block A (rows: the AOs of the segment's atoms j, first AO seg.y, `rows` of
// them; columns: the seg.w AOs of atom i from seg.x; column-major) at
// off[2k], then the mirror (rows: the AOs of atom i; columns: the first
// `mrows` AOs of the segment, those of the atoms j < i; column-major) at
// off[2k + 1]; off[2 nseg] is the unit's element count.
__host__ __device__ inline void tbx_elem_offsets(const int4 *seg, int nseg, int *off)
{
    int n = 0;
    for (int k = 0; k < nseg; ++k) {
        off[2 * k] = n;
        n += seg[k].w * (seg[k].z & 0xffff);
        off[2 * k + 1] = n;
        n += seg[k].w * (seg[k].z >> 16);
    }
    off[2 * nseg] = n;
}

// Element e of a forward unit (tbx_elem_offsets): its pair p (the unit's
// pairs run segment by segment, pext & 31, and within a segment by their
// first AO of atom j, so p is the last pair at or before (segment, AO)), the
// AO nul within atom i, the AO muj within atom j, and whether it lies in the
// mirror (rows of atom i, operator on j).
struct TbxElem {
    int p, nul, muj, mirror;
};
__host__ __device__ inline TbxElem tbx_elem_locate(const int *off, int nseg, const int4 *seg,
                                                    const TbxPair *pinf, const int *pext,
                                                    int npair, int e)
{
    int lo = 0, hi = 2 * nseg;  // last half-segment whose first element is <= e
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid;
    }
    const int k = lo >> 1;
    const int4 g = seg[k];
    const int local = e - off[lo];
    TbxElem r;
    r.mirror = lo & 1;
    int jao;  // molecule-local AO on the atom-j side
    if (!r.mirror) {
        const int rows = g.z & 0xffff;
        const int c = local / rows;
        r.nul = c;
        jao = g.y + (local - c * rows);
    } else {
        const int c = local / g.w;
        r.nul = local - c * g.w;
        jao = g.y + c;
    }
    int plo = 0, phi = npair;
    while (phi - plo > 1) {
        const int mid = (plo + phi) >> 1;
        const int sm = pext[mid] & 31;
        if (sm < k || (sm == k && pinf[mid].aoj <= jao)) plo = mid; else phi = mid;
    }
    r.p = plo;
    r.muj = jao - pinf[plo].aoj;
    return r;
}

// The pair search of tbx_elem_locate as a table, one byte per block-A row
// of the unit (at most 2 * TBX_UNIT_TASKS, since every row carries at least
// one element): roff[k] = the first entry of segment k (nseg + 1 of them),
// map[roff[k] + r] = the last pair of segment k whose first AO of atom j is
// at most the row's AO.  tbx_row_map_pair writes pair p's rows, from its
// first AO of atom j up to the next pair's of the segment (one pair per lane).
__host__ __device__ inline void tbx_row_offsets(const int4 *seg, int nseg, int *roff)
{
    int n = 0;
    for (int k = 0; k < nseg; ++k) {
        roff[k] = n;
        n += seg[k].z & 0xffff;
    }
    roff[nseg] = n;
}

__host__ __device__ inline void tbx_row_map_pair(const int4 *seg, const int *roff,
                                                 const TbxPair *pinf, const int *pext, int npair,
                                                 int p, unsigned char *map)
{
    const int k = pext[p] & 31;
    const int4 g = seg[k];
    const int rows = g.z & 0xffff;
    const int r1 = (p + 1 < npair && (pext[p + 1] & 31) == k) ? pinf[p + 1].aoj - g.y : rows;
    for (int r = pinf[p].aoj - g.y; r < r1; ++r)
        if (r >= 0 && r < rows) map[roff[k] + r] = (unsigned char)p;
}

// tbx_elem_locate with the pair read from the row map
__host__ __device__ inline TbxElem tbx_elem_locate_map(const int *off, int nseg, const int4 *seg,
                                                      const TbxPair *pinf, const int *roff,
                                                      const unsigned char *map, int e)
{
    int lo = 0, hi = 2 * nseg;  // last half-segment whose first element is <= e
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid;
    }
    const int k = lo >> 1;
    const int4 g = seg[k];
    const int local = e - off[lo];
    TbxElem r;
    r.mirror = lo & 1;
    int jao;  // molecule-local AO on the atom-j side
    if (!r.mirror) {
        const int rows = g.z & 0xffff;
        const int c = local / rows;
        r.nul = c;
        jao = g.y + (local - c * rows);
    } else {
        const int c = local / g.w;
        r.nul = local - c * g.w;
        jao = g.y + c;
    }
    r.p = map[roff[k] + (jao - g.y)];
    r.muj = jao - pinf[r.p].aoj;
    return r;
}

// Molecule / atom tables of a batch (device pointers).  "local" indices
// count from the start of the atom's molecule.
struct TbxMols {
    int nmol, nat;
    int ntrans;
    int zero_image;                // every structure has the one translation T = 0
    const double *trans;           // [3*ntrans] translation vectors of all molecules
    const int2 *mol_tr;            // [nmol] (first, count): translations of molecule m
    const int *mol_atom;           // [nmol+1] first atom
    const int *mol_nao;            // [nmol]
    const int *mol_cart;           // [nmol+1] first Cartesian function (global)
    const long long *mol_dense;    // [nmol+1] first dense element of S/H
    const int *at_sp;              // [nat]
    const int *at_mol;             // [nat]
    const int *at_sh;              // [nat] first shell (global) -> self-energies
    const int *at_ao;              // [nat] first AO, molecule-local
    const int *at_cart;            // [nat] first Cartesian function, molecule-local
    const int *cart_atom;          // [total Cartesian] atom (global)
    const int *ao_atom;            // [total AO] atom (global)
    const int *mol_ao;             // [nmol+1] first AO (global)
    int nunit;
    const TbxUnitHead *unit;       // [nunit] gradient units (j < i; j <= i with images); the
                                   // first nunit_near hold a pair within cutoff + skin of the
                                   // pair-order hint, the others none
    int nunit_near;
    const double *hint;            // [3*nat] hint geometry (nullptr: no hint)
    double hint_r2;                // squared displacement below which the far units stay dead
    const int *plan;               // their plans (16-byte words)
    int nfunit;
    const TbxUnitHead *funit;      // [nfunit] forward units (j <= i, on-site pairs included)
    int nfunit_far;                // forward units flagged TBX_UNIT_FAR
    const int *fplan;
    int nsh_tot, nao_tot;
    const int *sh_aog;             // [nsh_tot+1] first AO (global) of every shell
    int ntile;
    const int *tiles;              // [4*ntile] (mol, bi, bj, 0): 32 x 32 tile pairs bi <= bj
    int ndtile;
    const int *dtiles;             // [4*ndtile] the same for 64 x 64 (k_density)
                                   // of the dense blocks
};

// Work item of workgroup b of n: the hardware deals consecutive workgroups to
// the 8 XCDs of the MI355X in turn; this permutation gives each XCD a
// contiguous range of items instead (XCD x takes items [x per + min(x, rem),
// ...) with per = n / 8, rem = n % 8), so that items of one molecule share
// one L2.  A bijection of [0, n).
#define TBX_NXCD 8
__device__ __forceinline__ int tbx_xcd_block(int b, int n)
{
    const int x = b % TBX_NXCD, k = b / TBX_NXCD;
    const int per = n / TBX_NXCD, rem = n % TBX_NXCD;
    return x * per + min(x, rem) + k;
}
