/*
 * tbx.h -- C-ABI of the MI355X (gfx950) GFN2-xTB integral / zeroth-order
 * Hamiltonian path.  Plain pointers and sizes only; every array keeps the
 * Fortran column-major layout tblite uses, with 0-based indices.
 *
 * The entry points replace tblite's Fortran module procedures
 *   get_hamiltonian           src/tblite/xtb/h0.f90:220-429
 *   get_hamiltonian_gradient  src/tblite/xtb/h0.f90:432-666
 *   get_selfenergy            src/tblite/xtb/h0.f90:113-217
 * which have no C binding in the reference (they are called directly by
 * xtb_singlepoint, src/tblite/xtb/singlepoint.f90:228,241,370).  The
 * conventions follow tblite's own C API (the headers in include/tblite/): an opaque
 * context, int status returns (0 = success) and a retrievable message.
 *
 * Units: bohr and hartree.  Angular momenta l <= 2 (GFN2, xtb/gfn2.f90:928);
 * at most TBX_MAXSH shells per species and TBX_MAXG primitives per shell.
 */
#ifndef TBX_H
#define TBX_H

#ifdef __cplusplus
extern "C" {
#endif

#define TBX_MAXG 12  /* basis/type.f90:32 */
#define TBX_MAXSH 4  /* shells per species */
#define TBX_MAXL 2

/* status codes */
#define TBX_OK 0
#define TBX_ERR_ARG 1      /* invalid argument (null pointer, size, NaN) */
#define TBX_ERR_BASIS 2    /* unsupported basis (l > 2, too many shells/primitives) */
#define TBX_ERR_DEVICE 3   /* HIP runtime error */
#define TBX_ERR_ALLOC 4    /* device allocation failed */

typedef struct tbx_context tbx_context;

/* Species-resolved contracted Gaussian basis (basis_type, basis/type.f90:58-100).
 * Shell ish (0-based) of species isp lives in slot ish + msh*isp. */
typedef struct {
    int nid;              /* number of species */
    int msh;              /* leading dimension of the per-shell tables (<= TBX_MAXSH) */
    const int *nsh_id;    /* [nid]            shells per species */
    const int *ang;       /* [msh*nid]        angular momentum */
    const int *nprim;     /* [msh*nid]        primitives */
    const double *alpha;  /* [TBX_MAXG*msh*nid] exponents */
    const double *coeff;  /* [TBX_MAXG*msh*nid] contraction coefficients incl. normalisation */
    double intcut;        /* primitive screening threshold (25 for every accuracy) */
} tbx_basis;

/* Zeroth-order Hamiltonian parameters (tb_hamiltonian, xtb/h0.f90:39-67). */
typedef struct {
    const double *hscale; /* [msh*msh*nid*nid] hscale(jsh, ish, jsp, isp) */
    const double *shpoly; /* [msh*nid]         shpoly(ish, isp) */
    const double *rad;    /* [nid]             atomic radii (bohr) */
    /* Diatomic-frame scaling of the overlap in H0 (do_diat_scale with
     * ksig, kpi, kdel, xtb/h0.f90:64-65, 345-390; integral/diat_trafo.f90):
     * [nid*nid] each, ksig(izp, jzp) at izp + nid*jzp, or all three NULL
     * (off, as for GFN2).  With them, off-site H0 blocks use the overlap
     * scaled in the frame of the bond (sigma, pi, delta by ksig, kpi,
     * kdel); every species' shells must then be s, p, d in this order
     * (TBX_ERR_BASIS otherwise), and the gradient entry points return
     * TBX_ERR_ARG (the reference computes no gradient on its one scaled
     * path, the CEH guess, ceh/singlepoint.f90:166). */
    const double *ksig, *kpi, *kdel;
} tbx_h0par;

/* Structure and lattice images (structure_type + trans, cutoff.f90:97-121). */
typedef struct {
    int nat;
    const int *id;        /* [nat]      0-based species */
    const double *xyz;    /* [3*nat]    xyz(3, nat) */
    int ntrans;
    const double *trans;  /* [3*ntrans] translation vectors; one zero vector for molecules */
    double cutoff;        /* real-space cutoff, get_cutoff(bas) (basis/type.f90:227-246) */
} tbx_geom;

/* Half neighbour list in CSR form (adjacency_list, adjlist.f90:44-56), 0-based. */
typedef struct {
    const int *inl;   /* [nat] offset */
    const int *nnl;   /* [nat] count */
    int nnbr;         /* total entries */
    const int *nlat;  /* [nnbr] neighbouring atom (jat <= iat) */
    const int *nltr;  /* [nnbr] translation index */
} tbx_list;

/* --- context ----------------------------------------------------------- */

/* Create a context on HIP device `device`.  Returns NULL on failure. */
tbx_context *tbx_context_new(int device);
void tbx_context_delete(tbx_context *ctx);
/* Limit on the device memory the library allocates for this context and
 * the batches created on it (bytes; < 0: no limit, the default).  A call that
 * would need more fails with TBX_ERR_ALLOC before it changes any output or
 * buffer, and the context stays usable; no counterpart in the reference,
 * whose routines cannot fail (its errors go through the context,
 * context/type.F90). */
int tbx_context_set_memory_limit(tbx_context *ctx, long long bytes);
/* Device memory currently held by the context and its batches (bytes). */
long long tbx_context_memory_used(const tbx_context *ctx);
/* Message of the last failing call on this thread ("" if none). */
const char *tbx_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int tbx_version(void);

/* --- host-pointer entry points (drop-in for the Fortran procedures) ------ */

/* get_selfenergy (xtb/h0.f90:113-217), CN branch: se = selfenergy_sp - kcn*cn,
 * dsedcn = -kcn.  selfenergy_sp, kcn: [msh*nid]; cn may be NULL (no shift);
 * dsedcn may be NULL.  Host-side, no device work. */
int tbx_get_selfenergy(int nat, const int *id, const int *ish_at, const int *nsh_id, int msh,
                       const double *selfenergy_sp, const double *kcn, const double *cn,
                       double *selfenergy, double *dsedcn);

/* get_lattice_points_cutoff (cutoff.f90:97-121): the lattice points within
 * rthr of the cell lattice(3,3) (column k = lattice vector k, bohr), origin
 * first, in the reference's order; one zero vector when no direction of
 * periodic[3] is set.  On entry *ntrans is the room in trans [3 * *ntrans];
 * on return the number of points.  Too little room: TBX_ERR_ARG, with
 * *ntrans the number needed (call with *ntrans = 0 to ask).  Host only. */
int tbx_get_lattice_points(const int *periodic, const double *lattice, double rthr, int *ntrans,
                           double *trans);

/* get_hamiltonian (xtb/h0.f90:220-429): dense S(nao,nao), D(3,nao,nao),
 * Q(6,nao,nao) (traceless, operator on the column function) and H0(nao,nao),
 * all intent(out).  `list` may be NULL: neighbours are the pairs within
 * geom->cutoff, the rule new_adjacency_list applies (adjlist.f90:125-129).
 * iao_sh/ish_at: [nsh]/[nat] AO and shell offsets; selfenergy: [nsh]. */
int tbx_get_hamiltonian(tbx_context *ctx, const tbx_geom *geom, const tbx_list *list,
                        const tbx_basis *bas, const tbx_h0par *h0, const int *ish_at,
                        const int *iao_sh, int nao, const double *selfenergy, double *overlap,
                        double *dpint, double *qpint, double *hamiltonian);

/* The full double-loop integral builders get_multipole_integrals_lat,
 * get_dipole_integrals_lat and get_overlap_lat
 * (integral/native/integrals.f90:1394-1465, 994-1059, 411-470): dense
 * S(nao,nao), D(3,nao,nao), Q(6,nao,nao) (traceless, operator on the column
 * function), intent(out), over every atom pair and translation within
 * geom->cutoff.  The values are get_hamiltonian's S, D, Q.  Coincident
 * distinct atoms (r^2 below machine epsilon) contribute nothing, as in the
 * half neighbour list; the reference's double loop would add their blocks. */
int tbx_get_multipole_integrals(tbx_context *ctx, const tbx_geom *geom, const tbx_basis *bas,
                                const int *ish_at, const int *iao_sh, int nao, double *overlap,
                                double *dpint, double *qpint);
int tbx_get_dipole_integrals(tbx_context *ctx, const tbx_geom *geom, const tbx_basis *bas,
                             const int *ish_at, const int *iao_sh, int nao, double *overlap,
                             double *dpint);
int tbx_get_overlap(tbx_context *ctx, const tbx_geom *geom, const tbx_basis *bas,
                    const int *ish_at, const int *iao_sh, int nao, double *overlap);

/* get_hamiltonian_gradient (xtb/h0.f90:432-666).  pmat, xmat: (nao,nao,nspin)
 * (magnet representation, wavefunction/spin.f90:104-153); vao: (nao,nspin);
 * vdp: (3,nat,nspin); vqp: (6,nat,nspin).  dEdcn(nat), gradient(3,nat) and
 * sigma(3,3) are ACCUMULATED into, as the reference's intent(inout). */
int tbx_get_hamiltonian_gradient(tbx_context *ctx, const tbx_geom *geom, const tbx_list *list,
                                 const tbx_basis *bas, const tbx_h0par *h0, const int *ish_at,
                                 const int *iao_sh, int nao, const double *selfenergy,
                                 const double *dsedcn, int nspin, const double *vao,
                                 const double *vdp, const double *vqp, const double *pmat,
                                 const double *xmat, double *dEdcn, double *gradient,
                                 double *sigma);

/* Adjoint (backward) of get_hamiltonian (xtb/h0.f90:220-429), differentiated
 * the way get_hamiltonian_gradient (xtb/h0.f90:432-666) differentiates it.
 * Given adjoints sbar(nao,nao), hbar(nao,nao), dbar(3,nao,nao) and
 * qbar(6,nao,nao) of the four outputs, E = <sbar,S> + <hbar,H> + <dbar,D> +
 * <qbar,Q>; dE/dxyz into gradient(3,nat), dE/dselfenergy into dEdse(nsh)
 * and the strain derivative into sigma(3,3) (NULL: not wanted) are
 * ACCUMULATED into.  The integrals and their derivatives are recomputed, no
 * derivative tensor is stored.  The reference has no such routine: this is
 * the backward of the differentiable path (tbx_batch_adjoint) for one
 * structure, host arrays in and out, lattice images included. */
int tbx_get_hamiltonian_adjoint(tbx_context *ctx, const tbx_geom *geom, const tbx_list *list,
                                const tbx_basis *bas, const tbx_h0par *h0, const int *ish_at,
                                const int *iao_sh, int nao, const double *selfenergy,
                                const double *sbar, const double *hbar, const double *dbar,
                                const double *qbar, double *gradient, double *dEdse,
                                double *sigma);

/* --- device-resident batched path ---------------------------------------- */

/* A batch: nmol independent molecules sharing one species table.  Built once
 * (host tables uploaded), then run many times on device-resident inputs. */
typedef struct tbx_batch tbx_batch;

typedef struct {
    int nmol;
    const int *mol_nat;   /* [nmol] atoms per molecule */
    const int *id;        /* [sum nat] species, 0-based, into the basis tables */
    const tbx_basis *bas;
    const tbx_h0par *h0;
    double cutoff;        /* neighbour cutoff (bohr) */
} tbx_batch_desc;

/* Sizes of a batch: total atoms, shells, AOs and dense elements sum(nao^2). */
typedef struct {
    int nmol, nat, nsh, nao;
    long long ndense;
    long long npair; /* reserved, 0 */
} tbx_batch_sizes;

int tbx_batch_create(tbx_context *ctx, const tbx_batch_desc *desc, tbx_batch **out);
/* A batch of periodic structures (and molecules): molecule m has its own
 * mol_ntrans[m] >= 1 translation vectors, stored one molecule after the other
 * in trans [3 * sum mol_ntrans] (a molecule: one zero vector; a crystal: the
 * lattice points of get_lattice_points, cutoff.f90:97-121, e.g. from
 * tbx_get_lattice_points).  Pairs (i, j <= i, T) of every translation within
 * desc->cutoff enter the integrals, as new_adjacency_list puts them in the
 * half list (adjlist.f90:96-143); the coordination number sums over the same
 * translations within its own cutoff, so they should reach the larger of the
 * two.  The forward then adds the blocks of all images into zeroed outputs
 * with FP64 atomics. */
int tbx_batch_create_periodic(tbx_context *ctx, const tbx_batch_desc *desc,
                              const int *mol_ntrans, const double *trans, tbx_batch **out);
void tbx_batch_delete(tbx_batch *b);
int tbx_batch_sizes_get(const tbx_batch *b, tbx_batch_sizes *sizes);
/* Offsets into the packed per-molecule arrays: atom_off[nmol+1], sh_off[nmol+1],
 * ao_off[nmol+1], dense_off[nmol+1] (elements of S/H; x3 for D, x6 for Q). */
int tbx_batch_offsets(const tbx_batch *b, int *atom_off, int *sh_off, int *ao_off,
                      long long *dense_off);

/* Device pointers of the tbx_batch_* entry points must be ordinary
 * (coarse-grained) device memory, as hipMalloc and torch allocate it: the
 * gradient outputs (and the forward's with lattice images) are summed with
 * hardware FP64 atomics, which fine-grained memory (hipMallocManaged,
 * hipHostMalloc) does not guarantee.  Those sums run in no fixed order, so
 * dE/dR, dE/dse, dEdcn and sigma may differ in the last bits from run to run
 * (observed <= 1e-14 relative); every other output is bitwise reproducible. */

/* Forward: dense S, D, Q, H per molecule into device buffers of
 * ndense, 3*ndense, 6*ndense, ndense doubles.  d_xyz: [3*nat], d_se: [nsh],
 * device pointers.  Neighbours are the pairs within desc->cutoff, decided on
 * the device from d_xyz (no host neighbour list).  `stream` is the HIP
 * stream to launch on; NULL is the default (null) stream.  All tbx_batch_*
 * launches are asynchronous and allocate nothing on the device.  The dipole
 * and quadrupole outputs d_D, d_Q (and d_mpint of the SCF-loop consumers)
 * must be 16-byte aligned (hipMalloc and torch allocations are). */
int tbx_batch_forward(tbx_batch *b, const double *d_xyz, const double *d_se, double *d_S,
                      double *d_D, double *d_Q, double *d_H, void *stream);

/* Plain gradient contraction for every molecule (nspin = 1): d_P, d_W dense
 * per molecule (ndense each), d_vao [nao], d_vdp [3*nat], d_vqp [6*nat];
 * d_dsedcn [nsh].  d_dEdcn [nat], d_grad [3*nat] and d_sigma [9*nmol] are
 * OVERWRITTEN. */
int tbx_batch_gradient(tbx_batch *b, const double *d_xyz, const double *d_se,
                       const double *d_dsedcn, const double *d_P, const double *d_W,
                       const double *d_vao, const double *d_vdp, const double *d_vqp,
                       double *d_dEdcn, double *d_grad, double *d_sigma, void *stream);

/* tbx_batch_gradient given also the forward's own outputs d_S (ndense) and d_D
 * (3*ndense) at the same d_xyz and d_se, which an SCF caller holds from its
 * tbx_batch_forward (tblite keeps them in its integral container between
 * get_hamiltonian and get_hamiltonian_gradient).  As in
 * tbx_batch_adjoint_with_forward, the terms linear in S and D are contracted
 * per element from them; same outputs as tbx_batch_gradient, to rounding.
 * Structures with lattice images take tbx_batch_gradient's route.  No
 * counterpart in the reference. */
int tbx_batch_gradient_with_forward(tbx_batch *b, const double *d_xyz, const double *d_se,
                                    const double *d_S, const double *d_D,
                                    const double *d_dsedcn, const double *d_P,
                                    const double *d_W, const double *d_vao, const double *d_vdp,
                                    const double *d_vqp, double *d_dEdcn, double *d_grad,
                                    double *d_sigma, void *stream);

/* tbx_batch_gradient for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same arguments and results
 * as tbx_batch_gradient, but its dE/dse partial sums live in a static device
 * array of TBX_NOALLOC_MAX_SHELLS doubles (8 MiB, held by every device the
 * library runs on until the process ends) instead of the batch's own
 * buffer; it calls no hipMalloc, hipFree or allocating library routine.
 * Batches with more shells are rejected with TBX_ERR_ARG.  Every call on one
 * device shares that array: calls that overlap on one device (different
 * streams, or batches used from several threads) race. */
#define TBX_NOALLOC_MAX_SHELLS (1 << 20)
int tbx_batch_gradient_noalloc(tbx_batch *b, const double *d_xyz, const double *d_se,
                               const double *d_dsedcn, const double *d_P, const double *d_W,
                               const double *d_vao, const double *d_vdp, const double *d_vqp,
                               double *d_dEdcn, double *d_grad, double *d_sigma, void *stream);

/* Adjoint (backward) of tbx_batch_forward: given dense adjoints d_Sbar,
 * d_Hbar (ndense), d_Dbar (3*ndense), d_Qbar (6*ndense) of the forward
 * outputs, write dE/dxyz [3*nat] and dE/dse [nsh] (OVERWRITTEN).  The
 * integrals are recomputed, no derivative tensors are stored. */
int tbx_batch_adjoint(tbx_batch *b, const double *d_xyz, const double *d_se,
                      const double *d_Sbar, const double *d_Dbar, const double *d_Qbar,
                      const double *d_Hbar, double *d_dxyz, double *d_dse, void *stream);

/* tbx_batch_adjoint given also the forward's own outputs d_S (ndense) and d_D
 * (3*ndense) at the same d_xyz and d_se, as an autograd backward holds them.
 * The adjoint's terms that depend on the pair vector through the shift of
 * the moments to atom j and through the H0 distance polynomial are linear in
 * S and D, so they are contracted element by element from these instead of
 * being recomputed per primitive pair, which makes the kernel's work items
 * smaller.  Same outputs as tbx_batch_adjoint, to rounding.  A batch with
 * lattice images (more than one translation per structure) takes
 * tbx_batch_adjoint's route and leaves d_S, d_D unread.  The reference has
 * no such routine: it serves the backward of the differentiable path. */
int tbx_batch_adjoint_with_forward(tbx_batch *b, const double *d_xyz, const double *d_se,
                                   const double *d_S, const double *d_D, const double *d_Sbar,
                                   const double *d_Dbar, const double *d_Qbar,
                                   const double *d_Hbar, double *d_dxyz, double *d_dse,
                                   void *stream);

/* --- coordination number (device) ----------------------------------------
 * The GFN2 coordination number, mctc-lib's double-exponential counting
 * function as tblite creates it (xtb/gfn2.f90:640) and evaluates it
 * (xtb/singlepoint.f90:224):
 *   cn(i) = sum_{j,T; 0 < r <= cutoff} 1/(1+exp(-ka (r0/r-1))) *
 *           1/(1+exp(-kb ((r0+rshift)/r-1))),  r0 = rcov(i) + rcov(j),
 * over the atoms of i's molecule and the molecule's translations.  GFN2:
 * ka = 10, kb = 20, rshift = 2 bohr, cutoff 25 bohr, rcov = 4/3 of the D3
 * covalent radii (bohr).  tbx_batch_set_ncoord stores the radii ([nid],
 * per species) and constants; it waits for the device. */
int tbx_batch_set_ncoord(tbx_batch *b, const double *rcov, double ka, double kb, double rshift,
                         double cutoff);
/* d_cn [nat] OVERWRITTEN. */
int tbx_batch_ncoord(tbx_batch *b, const double *d_xyz, double *d_cn, void *stream);
/* Chain rule of the coordination number, what xtb/singlepoint.f90:375-378
 * does with dcndr and dcndL: d_grad [3*nat] += dcn/dxyz^T d_dEdcn and
 * d_sigma [9*nmol] += dcn/dstrain^T d_dEdcn per molecule (ACCUMULATED;
 * FP64 atomics for sigma). */
int tbx_batch_ncoord_gradient(tbx_batch *b, const double *d_xyz, const double *d_dEdcn,
                              double *d_grad, double *d_sigma, void *stream);

/* --- SCF-loop consumers of the resident integrals (batch layout) ---------
 * Dense spin arrays (d_P, d_h1) hold nspin (1 or 2) blocks of ndense
 * doubles; d_vao is [nao*nspin], d_vmp [vdim*nat*nspin].  The multipole
 * integrals are the forward's D (mpdim 3) or Q (mpdim 6) output, operator
 * on the column atom.  Asynchronous on `stream`, no allocation. */

/* add_vao_to_h1 (scf/potential.f90:192-219): d_h1 -= S (vao_r + vao_c)/2. */
int tbx_batch_add_vao_to_h1(tbx_batch *b, int nspin, const double *d_S, const double *d_vao,
                            double *d_h1, void *stream);
/* add_vmp_to_h1 (scf/potential.f90:221-255): the first min(mpdim, vdim)
 * components of d_vmp against d_mpint [mpdim*ndense]. */
int tbx_batch_add_vmp_to_h1(tbx_batch *b, int nspin, int mpdim, const double *d_mpint, int vdim,
                            const double *d_vmp, double *d_h1, void *stream);
/* get_mulliken_shell_charges (wavefunction/mulliken.f90:37-62): d_P in the
 * up/down representation; d_qsh [nsh*nspin] OVERWRITTEN in the magnet
 * representation (updown_to_magnet, wavefunction/spin.f90), + d_n0sh [nsh]
 * on the first. */
int tbx_batch_mulliken_shell_charges(tbx_batch *b, int nspin, const double *d_S,
                                     const double *d_P, const double *d_n0sh, double *d_qsh,
                                     void *stream);
/* get_mulliken_atomic_multipoles (wavefunction/mulliken.f90:65-98):
 * d_mpat [mpdim*nat*nspin] OVERWRITTEN (magnet representation). */
int tbx_batch_mulliken_atomic_multipoles(tbx_batch *b, int nspin, int mpdim,
                                         const double *d_mpint, const double *d_P,
                                         double *d_mpat, void *stream);
/* Milliseconds of the last add_vao_to_h1, add_vmp_to_h1, shell-charge and
 * atomic-multipole launches (ms[4]; -1 if not run), no reference counterpart. */
int tbx_batch_last_scf_timing(const tbx_batch *b, double *ms);

/* --- density matrices of the SCF (batch layout) ---------------------------
 * Per molecule, coefficients and densities are dense (nao_m x nao_m,
 * column-major, at dense_off[m]; column k of C is orbital k), orbital
 * energies and occupations [nao] at ao_off[m].  One spin channel per call
 * (the reference loops over spins, scf/diag.f90:62-108). */
/* sygvd (lapack/sygvd.f90, scf/diag.f90:84): H C = S C diag(eval) for every
 * molecule through rocSOLVER's dsygvd; d_H is overwritten with the
 * S-orthonormal eigenvectors C, d_eval [nao] with the ascending eigenvalues;
 * d_S is left unchanged (solved on a copy the batch holds).  Synchronises
 * the stream; TBX_ERR_ARG names a molecule whose S is not positive
 * definite.  rocSOLVER's own workspace is allocated by rocBLAS and is not
 * counted against tbx_context_set_memory_limit. */
int tbx_batch_sygvd(tbx_batch *b, double *d_H, const double *d_S, double *d_eval, void *stream);
/* get_fermi_filling (wavefunction/fermi.f90:31-117) for every molecule:
 * d_nel [nmol] electrons of this spin channel, kt (hartree), d_emo [nao]
 * ascending orbital energies; d_focc [nao] and d_efermi [nmol] OVERWRITTEN
 * (the reference's e_fermi: half the converged Fermi level). */
int tbx_batch_fermi_filling(tbx_batch *b, const double *d_nel, double kt, const double *d_emo,
                            double *d_focc, double *d_efermi, void *stream);
/* get_density_matrix (scf/diag.f90:150-166): d_P = C diag(focc) C^T, and the
 * energy-weighted form of get_wdensity (scf/diag.f90:110-143):
 * d_W = C diag(focc eval) C^T; FP64 matrix-core (MFMA) products, d_P / d_W
 * OVERWRITTEN. */
int tbx_batch_get_density_matrix(tbx_batch *b, const double *d_coeff, const double *d_focc,
                                 double *d_P, void *stream);
int tbx_batch_get_wdensity(tbx_batch *b, const double *d_coeff, const double *d_focc,
                           const double *d_eval, double *d_W, void *stream);

/* Kernel timing of the last batched call on the context's internal events:
 * milliseconds of the forward / pair-list / gradient / adjoint kernels. */
int tbx_batch_last_timing(const tbx_batch *b, double *ms_forward, double *ms_pairs,
                          double *ms_gradient, double *ms_adjoint);

/* Tile shapes of the batched kernels, no counterpart in the reference: the
 * half list of every molecule is cut into work units of at most
 * *_unit_tasks Cartesian-pair tasks (>= the largest atom pair; <= 256) and
 * *_unit_pairs atom pairs (<= 128); kept primitive pairs are processed in chunks of *_chunk
 * (1..64) per work item.  The pairs go into units in bands of *_band
 * consecutive row atoms (1..16), walked column by column, so that a unit
 * covers a few rows i and a run of columns j.  Defaults: forward 256, 128,
 * chunk 3, band 1; gradient 256, 128, chunk 1, band 4.
 * tbx_batch_set_tiles rebuilds the units (host) and waits for the whole
 * device (every stream) before it replaces them, so no launch still reads
 * the old units; tblite_amd/tune.py chooses them by timing the kernels.  A
 * HIP graph captured before a tbx_batch_set_tiles or tbx_batch_order_pairs
 * holds the old unit tables and counts: capture it again afterwards. */
typedef struct {
    int fwd_unit_tasks, fwd_unit_pairs, fwd_chunk;
    int grad_unit_tasks, grad_unit_pairs, grad_chunk;
    int fwd_band, grad_band;
} tbx_tiles;

int tbx_batch_set_tiles(tbx_batch *b, const tbx_tiles *tiles);
int tbx_batch_get_tiles(const tbx_batch *b, tbx_tiles *tiles);

/* Order of the atom pairs in the work units, no counterpart in the reference:
 * with a geometry hint xyz (host, bohr, [3*nat]), each molecule's pairs
 * closer than cutoff + skin there go into units first and the others into
 * units of their own, so that the gradient kernels skip whole units with no
 * pair inside the cutoff.  Every pair stays in exactly one unit and every
 * pair is still screened on the device, so the results do not depend on the
 * hint, only the speed does.  xyz NULL restores the geometry-independent
 * order.  Rebuilds the units (host) and waits for the whole device (every
 * stream) before it replaces them. */
int tbx_batch_order_pairs(tbx_batch *b, const double *xyz, double skin);

/* Work-unit plans of a batch description, built on the host only (no
 * context, no device) with `tiles` (NULL: the defaults), and checked against
 * what the kernels assume; no counterpart in the reference (an inspection
 * entry for tests and tooling).  Checked: every atom pair (i, j <= i) of every
 * molecule is in exactly one forward unit and every (i, j < i) in exactly one
 * gradient unit; every dense element of every molecule is written by exactly
 * one forward element as the kernel locates it from the unit's segment words,
 * inside its molecule's block, and reads only its unit's own Cartesian-pair
 * tasks; no unit exceeds the tiles' caps. */
typedef struct {
    long long nunit_fwd, nunit_grad;     /* work units */
    long long npair_fwd, npair_grad;     /* atom pairs over all units */
    long long ntask_fwd, ntask_grad;     /* Cartesian-pair tasks over all units */
    long long nelem;                     /* forward elements located */
    long long ndense;                    /* dense elements of the batch, sum of nao^2 */
    long long plan_bytes_fwd, plan_bytes_grad;
    long long bad_pairs;  /* pairs missing, repeated, or not of one molecule with j <= i */
    long long bad_elems;  /* elements written 0 or > 1 times, or reading outside the unit */
    long long bad_units;  /* units over a cap */
    long long nunit_grad_live; /* gradient units with a pair within the cutoff at xyz */
} tbx_plan_stats;

/* xyz, skin: the pair-order hint of tbx_batch_order_pairs (xyz may be NULL). */
int tbx_batch_plan_stats(const tbx_batch_desc *desc, const tbx_tiles *tiles, const double *xyz,
                         double skin, tbx_plan_stats *out);

/* Measurement helper, no counterpart in the reference: copies n doubles
 * d_src -> d_dst with a plain streaming kernel (k_copy) on `stream`.  Its
 * hardware-counter readings (FETCH_SIZE, WRITE_SIZE) for a known byte count
 * calibrate the HBM traffic reported for the integral kernels. */
int tbx_probe_copy(const double *d_src, double *d_dst, long long n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TBX_H */
