// tbx_capi.hip -- host implementation of the C-ABI declared in include/tbx.h.
//
// The host-pointer entry points (tbx_get_hamiltonian, tbx_get_hamiltonian_
// gradient) wrap a one-molecule batch: they upload the caller's arrays into
// device buffers cached on the context, launch the same kernels as the
// batched path on the context's stream, and copy the results back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <climits>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "../../include/tbx.h"
#include "tbx_tables.h"

// Default row atoms per band of the unit builders (build_units_list, the
// tiles' *_band).  Bands of 4 rows shorten the gradient's reads of the
// mirrored blocks (2048-molecule batch: adjoint given S, D 5.78 -> 5.66 ms;
// bands of 2 and 8: 5.67 and 5.76); the forward is fastest with single rows
// (bands of 2, 4, 8: 6.43, 6.47, 6.54 ms)
#define TBX_GRAD_BAND 4
#define TBX_FWD_BAND 1
#define TBX_MAX_BAND 16

extern "C" hipError_t tbx_launch_forward(const TbxSpecies *sp, const TbxMols *mol,
                                         const double *xyz, const double *se, double *S, double *D,
                                         double *Q, double *H, long long ndense, const int *nnl,
                                         int *flag, int chunk, hipStream_t stream);
extern "C" hipError_t tbx_launch_gradient(int mode, const TbxSpecies *sp, const TbxMols *mol,
                                          const double *xyz, const double *se,
                                          const double *P, const double *W, const double *vao,
                                          const double *vdp, const double *vqp, int nspin,
                                          long long pstride, long long vstride,
                                          const double *Sb, const double *Db, const double *Qb,
                                          const double *Hb, const double *Sf, const double *Df,
                                          double *grad, double *dse, double *sigma,
                                          const int *nnl, int *flag, int chunk,
                                          hipStream_t stream);
extern "C" hipError_t tbx_launch_dedcn(const TbxSpecies *sp, const TbxMols *mol,
                                       const double *dse, const double *dsedcn, const double *P,
                                       double *dEdcn, hipStream_t stream);
extern "C" hipError_t tbx_launch_fill(double *p, long long n, double v, hipStream_t stream);
extern "C" hipError_t tbx_launch_scf(int what, const TbxMols *mol, int nspin, long long dstride,
                                     int mpdim, int vdim, const double *S, const double *mp,
                                     const double *P, const double *vao, const double *vmp,
                                     const double *n0sh, double *h1, double *out,
                                     hipStream_t stream);
extern "C" hipError_t tbx_launch_copy(const double *src, double *dst, long long n,
                                      hipStream_t stream);
extern "C" hipError_t tbx_launch_density(const TbxMols *mol, const double *C, const double *focc,
                                         const double *eval, double *P, hipStream_t stream);
extern "C" hipError_t tbx_launch_fermi(const TbxMols *mol, const double *nel, double kt,
                                       const double *emo, double *occ, double *efermi,
                                       hipStream_t stream);
extern "C" hipError_t tbx_launch_ncoord(const TbxMols *mol, const double *rcov, double ka,
                                        double kb, double rshift, double cutoff,
                                        const double *xyz, double *cn, hipStream_t stream);
extern "C" hipError_t tbx_launch_ncoord_grad(const TbxMols *mol, const double *rcov, double ka,
                                             double kb, double rshift, double cutoff,
                                             const double *xyz, const double *dEdcn,
                                             double *grad, double *sigma, hipStream_t stream);

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? TBX_ERR_ALLOC : TBX_ERR_DEVICE,     \
                        "%s: %s", #expr, hipGetErrorString(e_));                        \
    } while (0)

// Device memory the library holds for one context and its batches, and the
// limit a caller may set on it (tbx_context_set_memory_limit).
struct Budget {
    size_t limit = SIZE_MAX, used = 0;
};

// Device buffer that grows on demand and is freed with its owner.  A request
// beyond the owner's budget fails with hipErrorOutOfMemory and leaves the
// buffer as it was.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    Budget *budget = nullptr;
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        const size_t want = std::max<size_t>(bytes, 256);
        if (budget && (budget->used - cap > budget->limit || want > budget->limit - (budget->used - cap)))
            return hipErrorOutOfMemory;
        if (p) (void)hipFree(p);
        if (budget) budget->used -= cap;
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) {
            cap = want;
            if (budget) budget->used += cap;
        } else {
            p = nullptr;
        }
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    void release()
    {
        if (p) (void)hipFree(p);
        if (budget) budget->used -= cap;
        p = nullptr;
        cap = 0;
    }
};

template <class T>
hipError_t upload(DevBuf &b, const std::vector<T> &v, hipStream_t s)
{
    hipError_t e = b.reserve(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s);
}

} // namespace

// ---------------------------------------------------------------------------
// Static device scratch of tbx_batch_gradient_noalloc: dE/dse per shell
// (TBX_NOALLOC_MAX_SHELLS doubles, 8 MiB; every device the library runs on
// holds one copy until the process ends).  In host code `scratch` is only a
// stand-in for the symbol: its device address, one per device, is looked up
// once with hipGetSymbolAddress when the first context on that device is
// created, and only that address is passed to kernels.
static __device__ __align__(256) double scratch[TBX_NOALLOC_MAX_SHELLS];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // by device

// Device address of `scratch` on `device` (which must be current).
int scratch_init(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return TBX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(TBX_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return TBX_OK;
}

double *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}
} // namespace

// ---------------------------------------------------------------------------
// host-side tables

struct HostSpecies {
    int nsp = 0, msh = 0, maxnc = 0;
    std::vector<int> nsh, ang, nprim, ncart, nao, cart_sh, cart_exp, ao_sh, ao_m, sh_cart, sh_ao,
        prim_off, pp_off, ao_info, ao_desc;
    std::vector<double> alpha, coeff, shpoly, rad, hscale, prim, pp, pp_mu, onsite;
    std::vector<double> kdiat;  // (ksig, kpi, kdel) at 3 * (isp + nsp * jsp); empty: off
    bool diat = false;
    double intcut = 25.0;
};

// The gradient kernels do not differentiate the diatomic-frame scaling (the
// reference's one scaled path, the CEH guess, ceh/singlepoint.f90:166, takes
// no gradient).
static const char kNoDiatGradient[] = "diatomic-frame scaling: no gradient of the scaled H0";

static int build_species(const tbx_basis *bas, const tbx_h0par *h0, HostSpecies &hs)
{
    if (!bas || !h0) return fail(TBX_ERR_ARG, "basis or h0 parameters missing");
    if (bas->nid <= 0 || bas->msh <= 0 || bas->msh > TBX_MAXSH)
        return fail(TBX_ERR_BASIS, "nid=%d msh=%d outside 1..%d", bas->nid, bas->msh, TBX_MAXSH);
    if (!bas->nsh_id || !bas->ang || !bas->nprim || !bas->alpha || !bas->coeff || !h0->hscale ||
        !h0->shpoly || !h0->rad)
        return fail(TBX_ERR_ARG, "null basis/h0 table");
    if (!(bas->intcut > 0.0) || !std::isfinite(bas->intcut))
        return fail(TBX_ERR_ARG, "invalid primitive screening threshold intcut=%g", bas->intcut);
    const int nsp = bas->nid, msh = bas->msh;
    hs.nsp = nsp;
    hs.msh = msh;
    hs.intcut = bas->intcut;
    hs.nsh.assign(bas->nsh_id, bas->nsh_id + nsp);
    hs.ang.assign(bas->ang, bas->ang + nsp * msh);
    hs.nprim.assign(bas->nprim, bas->nprim + nsp * msh);
    hs.alpha.assign(bas->alpha, bas->alpha + TBX_MAXG * nsp * msh);
    hs.coeff.assign(bas->coeff, bas->coeff + TBX_MAXG * nsp * msh);
    hs.shpoly.assign(h0->shpoly, h0->shpoly + nsp * msh);
    hs.rad.assign(h0->rad, h0->rad + nsp);
    hs.hscale.assign(h0->hscale, h0->hscale + (size_t)msh * msh * nsp * nsp);
    // diatomic-frame scaling (xtb/h0.f90:64-65, 107-108): all three tables or none
    const int nk = (h0->ksig != nullptr) + (h0->kpi != nullptr) + (h0->kdel != nullptr);
    if (nk != 0 && nk != 3) return fail(TBX_ERR_ARG, "ksig, kpi and kdel must be given together");
    hs.diat = nk == 3;
    hs.kdiat.clear();
    if (hs.diat) {
        hs.kdiat.resize((size_t)3 * nsp * nsp);
        for (size_t k = 0; k < (size_t)nsp * nsp; ++k) {
            const double v[3] = {h0->ksig[k], h0->kpi[k], h0->kdel[k]};
            for (int c = 0; c < 3; ++c) {
                if (!std::isfinite(v[c]))
                    return fail(TBX_ERR_ARG, "non-finite diatomic scaling factor (pair %zu)", k);
                hs.kdiat[3 * k + c] = v[c];
            }
        }
        // the frame slots of an atom are its shells s, p, d in this order
        // (smap(ish - 1), xtb/h0.f90:288-294)
        for (int isp = 0; isp < nsp; ++isp)
            for (int ish = 0; ish < std::min(bas->nsh_id[isp], msh); ++ish)
                if (bas->ang[ish + msh * isp] != ish)
                    return fail(TBX_ERR_BASIS,
                                "diatomic-frame scaling needs the shells of every species in the "
                                "order s, p, d (species %d shell %d has l=%d)",
                                isp, ish, bas->ang[ish + msh * isp]);
    }
    hs.ncart.assign(nsp, 0);
    hs.nao.assign(nsp, 0);
    hs.cart_sh.assign((size_t)nsp * TBX_MAXCART, 0);
    hs.cart_exp.assign((size_t)nsp * TBX_MAXCART, 0);
    hs.ao_sh.assign((size_t)nsp * TBX_MAXAO, 0);
    hs.ao_m.assign((size_t)nsp * TBX_MAXAO, 0);
    hs.sh_cart.assign((size_t)nsp * msh, 0);
    hs.sh_ao.assign((size_t)nsp * msh, 0);
    for (int isp = 0; isp < nsp; ++isp) {
        if (hs.nsh[isp] < 1 || hs.nsh[isp] > msh)
            return fail(TBX_ERR_BASIS, "species %d has %d shells (msh=%d)", isp, hs.nsh[isp], msh);
        int nc = 0, na = 0;
        for (int ish = 0; ish < hs.nsh[isp]; ++ish) {
            const int slot = ish + msh * isp;
            const int l = hs.ang[slot];
            if (l < 0 || l > TBX_MAXL)
                return fail(TBX_ERR_BASIS, "angular momentum l=%d > %d (species %d shell %d)", l,
                            TBX_MAXL, isp, ish);
            if (hs.nprim[slot] < 1 || hs.nprim[slot] > TBX_MAXG)
                return fail(TBX_ERR_BASIS, "nprim=%d outside 1..%d", hs.nprim[slot], TBX_MAXG);
            for (int p = 0; p < hs.nprim[slot]; ++p)
                if (!(hs.alpha[TBX_MAXG * slot + p] > 0.0) ||
                    !std::isfinite(hs.coeff[TBX_MAXG * slot + p]))
                    return fail(TBX_ERR_BASIS, "invalid primitive in species %d shell %d", isp, ish);
            hs.sh_cart[slot] = nc;
            hs.sh_ao[slot] = na;
            int k = 0;  // CCA order: decreasing lx, then decreasing ly
            for (int lx = l; lx >= 0; --lx)
                for (int ly = l - lx; ly >= 0; --ly, ++k) {
                    const int lz = l - lx - ly;
                    if (nc >= TBX_MAXCART) return fail(TBX_ERR_BASIS, "too many Cartesian functions");
                    hs.cart_sh[isp * TBX_MAXCART + nc] = ish;
                    hs.cart_exp[isp * TBX_MAXCART + nc] = lx | (ly << 4) | (lz << 8) | (k << 12);
                    ++nc;
                }
            for (int mm = 0; mm < 2 * l + 1; ++mm) {
                if (na >= TBX_MAXAO) return fail(TBX_ERR_BASIS, "too many spherical functions");
                hs.ao_sh[isp * TBX_MAXAO + na] = ish;
                hs.ao_m[isp * TBX_MAXAO + na] = mm;
                ++na;
            }
        }
        hs.ncart[isp] = nc;
        hs.nao[isp] = na;
        hs.maxnc = std::max(hs.maxnc, nc);
    }
    // the kernels' windows hold at least one atom: (tasks per window)/nc >= nc
    if (hs.maxnc > 16) return fail(TBX_ERR_BASIS, "%d Cartesian functions per atom (max 16)", hs.maxnc);
    // packed primitive list, staged whole in LDS by the gradient kernels
    hs.prim_off.assign((size_t)nsp * msh, 0);
    hs.prim.clear();
    for (int isp = 0; isp < nsp; ++isp)
        for (int ish = 0; ish < hs.nsh[isp]; ++ish) {
            const int slot = ish + msh * isp;
            hs.prim_off[slot] = (int)(hs.prim.size() / 2);
            for (int p = 0; p < hs.nprim[slot]; ++p) {
                hs.prim.push_back(hs.alpha[TBX_MAXG * slot + p]);
                hs.prim.push_back(hs.coeff[TBX_MAXG * slot + p]);
            }
        }
    if (hs.prim.size() / 2 > TBX_MAXPRIM)
        return fail(TBX_ERR_BASIS, "%zu primitives over all species (max %d)", hs.prim.size() / 2,
                    TBX_MAXPRIM);
    // primitive pairs of every slot pair, sorted by the reduced exponent so
    // that the pairs kept by the screening ai aj r2/(ai+aj) <= intcut form a prefix
    const int nslot = nsp * msh;
    hs.pp_off.assign((size_t)nslot * nslot + 1, 0);
    hs.pp.clear();
    std::vector<std::array<double, TBX_PP_N>> ent;
    for (int si = 0; si < nslot; ++si)
        for (int sj = 0; sj < nslot; ++sj) {
            hs.pp_off[(size_t)si * nslot + sj] = (int)(hs.pp.size() / TBX_PP_N);
            const bool vi = (si % msh) < hs.nsh[si / msh], vj = (sj % msh) < hs.nsh[sj / msh];
            if (!vi || !vj) continue;
            ent.clear();
            for (int ip = 0; ip < hs.nprim[si]; ++ip)
                for (int jp = 0; jp < hs.nprim[sj]; ++jp) {
                    const double ai = hs.alpha[TBX_MAXG * si + ip], aj = hs.alpha[TBX_MAXG * sj + jp];
                    const double oab = 1.0 / (ai + aj);
                    std::array<double, TBX_PP_N> e;
                    e[TBX_PP_MU] = ai * aj * oab;
                    e[TBX_PP_AJO] = aj * oab;
                    e[TBX_PP_AIO] = ai * oab;
                    e[TBX_PP_H] = 0.5 * oab;
                    e[TBX_PP_AI] = ai;
                    e[TBX_PP_PRE] = 5.5683279968317078453 * oab * std::sqrt(oab) *
                                    hs.coeff[TBX_MAXG * si + ip] * hs.coeff[TBX_MAXG * sj + jp];
                    ent.push_back(e);
                }
            std::stable_sort(ent.begin(), ent.end(),
                             [](const std::array<double, TBX_PP_N> &a,
                                const std::array<double, TBX_PP_N> &b) {
                                 return a[TBX_PP_MU] < b[TBX_PP_MU];
                             });
            for (const auto &e : ent) hs.pp.insert(hs.pp.end(), e.begin(), e.end());
        }
    hs.pp_off[(size_t)nslot * nslot] = (int)(hs.pp.size() / TBX_PP_N);
    hs.pp_mu.resize(hs.pp.size() / TBX_PP_N);
    for (size_t k = 0; k < hs.pp_mu.size(); ++k) hs.pp_mu[k] = hs.pp[TBX_PP_N * k + TBX_PP_MU];
    // same-centre moments: sums over the primitive pairs of pre * prod_k
    // <y^(a_k + b_k + m_k)>, with the even moments (2n-1)!! h^n of the product
    // Gaussian; m = 0 (S), one unit vector (D), two (Q: xx xy yy xz yz zz)
    auto moment = [](double h, int t) {
        switch (t) {
        case 0: return 1.0;
        case 2: return h;
        case 4: return 3.0 * h * h;
        case 6: return 15.0 * h * h * h;
        case 8: return 105.0 * h * h * h * h;
        default: return 0.0;
        }
    };
    static const int kMom[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                    {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};
    hs.onsite.assign((size_t)nsp * TBX_MAXCART * TBX_MAXCART * 10, 0.0);
    for (int isp = 0; isp < nsp; ++isp)
        for (int ci = 0; ci < hs.ncart[isp]; ++ci)
            for (int cj = 0; cj < hs.ncart[isp]; ++cj) {
                const int ei = hs.cart_exp[isp * TBX_MAXCART + ci];
                const int ej = hs.cart_exp[isp * TBX_MAXCART + cj];
                const int si = hs.cart_sh[isp * TBX_MAXCART + ci] + msh * isp;
                const int sj = hs.cart_sh[isp * TBX_MAXCART + cj] + msh * isp;
                const int e[3] = {(ei & 15) + (ej & 15), ((ei >> 4) & 15) + ((ej >> 4) & 15),
                                  ((ei >> 8) & 15) + ((ej >> 8) & 15)};
                double *out = &hs.onsite[(((size_t)isp * TBX_MAXCART + ci) * TBX_MAXCART + cj) * 10];
                for (int q = hs.pp_off[(size_t)si * nslot + sj];
                     q < hs.pp_off[(size_t)si * nslot + sj + 1]; ++q) {
                    const double h = hs.pp[TBX_PP_N * q + TBX_PP_H];
                    const double pre = hs.pp[TBX_PP_N * q + TBX_PP_PRE];
                    for (int k = 0; k < 10; ++k)
                        out[k] += pre * moment(h, e[0] + kMom[k][0]) *
                                  moment(h, e[1] + kMom[k][1]) * moment(h, e[2] + kMom[k][2]);
                }
            }
    hs.ao_info.assign((size_t)nsp * TBX_MAXAO, 0);
    for (size_t k = 0; k < hs.ao_info.size(); ++k) {
        const int isp = (int)(k / TBX_MAXAO);
        const int sh = hs.ao_sh[k];
        hs.ao_info[k] = sh | (hs.ao_m[k] << 4) | (hs.ang[sh + msh * isp] << 8);
    }
    hs.ao_desc.assign(nsp, 0);
    for (int isp = 0; isp < nsp; ++isp) {
        int d = hs.nsh[isp];
        for (int ish = 0; ish < hs.nsh[isp]; ++ish) d |= hs.ang[ish + msh * isp] << (3 + 2 * ish);
        hs.ao_desc[isp] = d;
    }
    return TBX_OK;
}

struct HostMols {
    int nmol = 0, nat = 0, nsh = 0, nao = 0, ncart = 0;
    // translation vectors (3 per entry) and, per molecule, the (first, count)
    // of its own into that table: one zero vector for molecules
    std::vector<double> trans;
    std::vector<int2> mol_tr;
    long long ndense = 0;
    std::vector<int> mol_atom, mol_nao, mol_cart, mol_ao, at_sp, at_mol, at_sh, at_ao, at_cart,
        cart_atom, ao_atom, sh_off, unit_seg, funit_seg, uplan, fplan, sh_aog, tiles, dtiles;
    std::vector<long long> mol_dense;
    std::vector<TbxSeg> useg, fseg;
    std::vector<TbxUnitHead> uhead, fhead;
    // optional geometry hint (bohr, [3*nat]) for the order of the pairs in the
    // units: pairs within sqrt(near2) there come first in each molecule
    std::vector<double> hint_xyz;
    double near2 = 0.0, skin = 0.0;
    int nunit_near = 0;  // gradient units holding a pair within cutoff + skin at the hint
    int nfunit_far = 0;  // forward units holding none
};

// Work units (see TbxSeg): the half list of every molecule, cut into units of
// bounded task count.  Gradient: j < i (j <= i with lattice images); forward:
// j <= i, so that the on-site blocks are written by the same kernel.  With a
// geometry hint (molecules only), each molecule's pairs within the hint
// distance come first and the others follow in units of their own: every pair
// is still in exactly one unit and the kernels still screen every pair, so
// the hint changes the speed, never the results.
static void build_units_list(const HostSpecies &hs, const HostMols &hm, bool diag,
                             int max_tasks, int max_pairs, int band, std::vector<TbxSeg> &useg,
                             std::vector<int> &unit_seg, std::vector<char> *unit_far = nullptr)
{
    band = std::max(band, 1);
    const bool hinted = hm.hint_xyz.size() == 3 * (size_t)hm.nat;
    auto near = [&](int i, int j, int itr) {
        if (!hinted) return true;
        double r2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = hm.hint_xyz[3 * i + c] - hm.hint_xyz[3 * j + c] - hm.trans[3 * itr + c];
            r2 += d * d;
        }
        return i == j && r2 < 1e-12 ? true : r2 <= hm.near2;
    };
    useg.clear();
    unit_seg.assign(1, 0);
    if (unit_far) unit_far->clear();
    int tasks = 0, pairs = 0, segs = 0, cur_pass = 0;
    bool open_seg = false;
    auto close_unit = [&]() {
        if (segs > 0) {
            unit_seg.push_back((int)useg.size());
            if (unit_far) unit_far->push_back(cur_pass == 1 ? 1 : 0);
        }
        tasks = pairs = segs = 0;
        open_seg = false;
    };
    // the open segment of each row of the band in the current unit (-1: none)
    std::vector<int> seg_of(band > 0 ? band : 1, -1);
    auto close_all = [&]() {
        close_unit();
        std::fill(seg_of.begin(), seg_of.end(), -1);
    };
    for (int m = 0; m < hm.nmol; ++m) {
        const int a0 = hm.mol_atom[m], a1 = hm.mol_atom[m + 1];
        const int2 tr = hm.mol_tr[m];
        for (int pass = 0; pass < (hinted ? 2 : 1); ++pass) {
            cur_pass = pass;
            for (int itr = tr.x; itr < tr.x + tr.y; ++itr)
                // bands of `band` row atoms, walked column by column: a unit's
                // blocks then span a few rows i as well as a run of columns j,
                // so the mirrored blocks (rows of the atoms i) are longer runs
                for (int i0 = a0; i0 < a1; i0 += band) {
                    const int i1 = std::min(i0 + band, a1);
                    // with translations the self images (i, i, T) are pairs too
                    const int jlast = (diag || tr.y > 1) ? i1 : i1 - 1;
                    std::fill(seg_of.begin(), seg_of.end(), -1);
                    for (int j = a0; j < jlast; ++j) {
                        const int ncj = hs.ncart[hm.at_sp[j]];
                        for (int i = i0; i < i1; ++i) {
                            int &sg = seg_of[i - i0];
                            const int jend = (diag || tr.y > 1) ? i + 1 : i;
                            if (j >= jend || (hinted && near(i, j, itr) != (pass == 0))) {
                                sg = -1;  // a segment is a run of consecutive j
                                continue;
                            }
                            const int c = hs.ncart[hm.at_sp[i]] * ncj;
                            if (tasks + c > max_tasks || pairs == max_pairs ||
                                (sg < 0 && segs == TBX_UNIT_SEGS))
                                close_all();
                            if (sg < 0) {
                                sg = (int)useg.size();
                                useg.push_back(TbxSeg{i, j, j, itr});
                                ++segs;
                            }
                            useg[sg].j1 = j + 1;
                            tasks += c;
                            ++pairs;
                        }
                    }
                }
            close_all();  // units never span molecules, nor near and far pairs
        }
    }
}

// Plans of a list of units (layout: TbxUnitHead).  The units respect the
// caps max_tasks, max_pairs (those of the kernel that runs them) and
// TBX_UNIT_SEGS by construction.  Returns false if a unit breaks the caps or
// its plan does not have the length the kernels read (load_unit).
static bool build_plans(const HostSpecies &hs, const HostMols &hm, const std::vector<TbxSeg> &seg,
                        const std::vector<int> &unit_seg, bool with_elems, int max_tasks,
                        int max_pairs, std::vector<TbxUnitHead> &head, std::vector<int> &plan)
{
    const int nunit = (int)unit_seg.size() - 1;
    head.assign(nunit, TbxUnitHead{});
    plan.clear();
    std::vector<TbxPair> pairs;
    std::vector<int> pext, sinfo, spl;
    std::vector<unsigned char> spidx;
    auto pad4 = [&plan]() {
        while (plan.size() % 4) plan.push_back(0);
    };
    for (int u = 0; u < nunit; ++u) {
        const int s0 = unit_seg[u], s1 = unit_seg[u + 1];
        pairs.clear();
        pext.clear();
        sinfo.clear();
        for (int k = s0; k < s1; ++k) {
            const TbxSeg &g = seg[k];
            const int first = (int)pairs.size();
            int rows = 0, mrows = 0;
            for (int j = g.j0; j < g.j1; ++j) {
                TbxPair p;
                p.i = g.i;
                p.j = j;
                p.spi = hm.at_sp[g.i];
                p.spj = hm.at_sp[j];
                p.aoi = hm.at_ao[g.i];
                p.aoj = hm.at_ao[j];
                p.shi = hm.at_sh[g.i];
                p.shj = hm.at_sh[j];
                p.di = hs.ao_desc[p.spi];
                p.dj = hs.ao_desc[p.spj];
                pairs.push_back(p);
                pext.push_back((g.itr << 5) | (k - s0));
                rows += hs.nao[p.spj];
                if (j < g.i) mrows += hs.nao[p.spj];  // mirrored blocks: j < i
            }
            const int spi = hm.at_sp[g.i];
            sinfo.insert(sinfo.end(), {first, hm.at_ao[g.j0], rows | (mrows << 16), hs.nao[spi]});
        }
        const int npair = (int)pairs.size(), nseg = s1 - s0;
        // shell pairs, type-major; within a type by pair, then ish, then jsh
        spl.clear();
        spidx.assign((size_t)16 * npair, 0);
        int toff[10], tsp[10], ntask = 0;
        for (int t = 0; t < 9; ++t) {
            toff[t] = ntask;
            tsp[t] = (int)spl.size();
            const int li = t / 3, lj = t % 3;
            for (int p = 0; p < npair; ++p) {
                const TbxPair &pr = pairs[p];
                for (int qi = 0; qi < hs.nsh[pr.spi]; ++qi) {
                    if (hs.ang[qi + hs.msh * pr.spi] != li) continue;
                    for (int qj = 0; qj < hs.nsh[pr.spj]; ++qj) {
                        if (hs.ang[qj + hs.msh * pr.spj] != lj) continue;
                        spidx[(size_t)16 * p + 4 * qi + qj] = (unsigned char)spl.size();
                        spl.push_back(p | (qj << 8) | (qi << 12) | (t << 16));
                        ntask += (li + 1) * (li + 2) / 2 * ((lj + 1) * (lj + 2) / 2);
                    }
                }
            }
        }
        toff[9] = ntask;
        tsp[9] = (int)spl.size();
        const int m = hm.at_mol[seg[s0].i];
        TbxUnitHead &h = head[u];
        h.dense0 = hm.mol_dense[m];
        h.plan = (int)(plan.size() / 4);
        h.counts = npair | (nseg << 8) | ((int)spl.size() << 16);
        h.nao = hm.mol_nao[m];
        h.ao_m = hm.mol_ao[m];
        h.mol = m;
        for (const TbxPair &pr : pairs)
            plan.insert(plan.end(), {pr.i, pr.j, pr.spi, pr.spj, pr.aoi, pr.aoj, pr.shi, pr.shj,
                                     pr.di, pr.dj});
        pad4();
        plan.insert(plan.end(), pext.begin(), pext.end());
        pad4();
        // the kernels read the segments' first AO of atom i where the host
        // keeps the first pair
        for (int k = 0; k < nseg; ++k)
            plan.insert(plan.end(), {hm.at_ao[seg[s0 + k].i], sinfo[4 * k + 1], sinfo[4 * k + 2],
                                     sinfo[4 * k + 3]});
        plan.insert(plan.end(), spl.begin(), spl.end());
        pad4();
        plan.insert(plan.end(), toff, toff + 10);
        plan.insert(plan.end(), tsp, tsp + 10);
        const int nspair = (int)spl.size();
        const long long words = (10 * npair + 3) / 4 + (npair + 3) / 4 + nseg + (nspair + 3) / 4 + 5;
        if (npair < 1 || npair > max_pairs || nseg > TBX_UNIT_SEGS || nspair > max_tasks ||
            ntask > max_tasks ||
            (long long)plan.size() != 4 * ((long long)h.plan + words))
            return false;
        h.nel = 0;
        if (with_elems) {
            // dense elements: block A and mirror of every segment (tbx_elem_offsets)
            for (int k = 0; k < nseg; ++k) {
                const int rows = sinfo[4 * k + 2] & 0xffff, mrows = sinfo[4 * k + 2] >> 16;
                h.nel += sinfo[4 * k + 3] * (rows + mrows);
            }
            if (h.nel > 2 * max_tasks) return false;  // two elements per lane
        }
        if (plan.size() / 4 > (size_t)INT_MAX) return false;
    }
    return true;
}

static int build_units(const HostSpecies &hs, HostMols &hm, const tbx_tiles &t)
{
    std::vector<char> far, ffar;
    build_units_list(hs, hm, false, t.grad_unit_tasks, t.grad_unit_pairs, t.grad_band, hm.useg,
                     hm.unit_seg, &far);
    build_units_list(hs, hm, true, t.fwd_unit_tasks, t.fwd_unit_pairs, t.fwd_band, hm.fseg,
                     hm.funit_seg, &ffar);
    if (!build_plans(hs, hm, hm.useg, hm.unit_seg, false, TBX_GRAD_UNIT_TASKS,
                     TBX_GRAD_UNIT_PAIRS, hm.uhead, hm.uplan) ||
        !build_plans(hs, hm, hm.fseg, hm.funit_seg, true, TBX_UNIT_TASKS, TBX_UNIT_PAIRS,
                     hm.fhead, hm.fplan))
        return fail(TBX_ERR_ARG, "work-unit plan outside the kernel limits");
    // gradient units of pairs beyond cutoff + skin at the hint go last: the
    // kernels skip them while every atom stays within skin/2 of the hint
    std::vector<TbxUnitHead> near, farh;
    for (size_t u = 0; u < hm.uhead.size(); ++u) (far[u] ? farh : near).push_back(hm.uhead[u]);
    hm.nunit_near = (int)near.size();
    near.insert(near.end(), farh.begin(), farh.end());
    hm.uhead.swap(near);
    // forward units of such pairs stay in place (the XCD-contiguous unit
    // order), flagged: while the hint holds they only write zero blocks
    hm.nfunit_far = 0;
    for (size_t u = 0; u < hm.fhead.size(); ++u)
        if (ffar[u]) {
            hm.fhead[u].nel |= TBX_UNIT_FAR;
            ++hm.nfunit_far;
        }
    return TBX_OK;
}

// defaults: the autotuner's choice on the 2048-molecule batch
static const tbx_tiles kDefaultTiles = {TBX_UNIT_TASKS,      TBX_UNIT_PAIRS,      3,
                                        TBX_GRAD_UNIT_TASKS, TBX_GRAD_UNIT_PAIRS, 1,
                                        TBX_FWD_BAND,        TBX_GRAD_BAND};

// Tile shapes must fit the kernels' LDS arrays and hold the largest atom pair.
static int check_tiles(const HostSpecies &hs, const tbx_tiles *t)
{
    const int need = hs.maxnc * hs.maxnc;
    const int tasks[2] = {t->fwd_unit_tasks, t->grad_unit_tasks};
    const int pairs[2] = {t->fwd_unit_pairs, t->grad_unit_pairs};
    const int chunk[2] = {t->fwd_chunk, t->grad_chunk};
    const int max_tasks[2] = {TBX_UNIT_TASKS, TBX_GRAD_UNIT_TASKS};
    const int max_pairs[2] = {TBX_UNIT_PAIRS, TBX_GRAD_UNIT_PAIRS};
    for (int k = 0; k < 2; ++k) {
        if (tasks[k] < need || tasks[k] > max_tasks[k])
            return fail(TBX_ERR_ARG, "unit tasks %d outside %d..%d", tasks[k], need,
                        max_tasks[k]);
        if (pairs[k] < 1 || pairs[k] > max_pairs[k])
            return fail(TBX_ERR_ARG, "unit pairs %d outside 1..%d", pairs[k], max_pairs[k]);
        if (chunk[k] < 1 || chunk[k] > 64)
            return fail(TBX_ERR_ARG, "chunk %d outside 1..64", chunk[k]);
    }
    for (int band : {t->fwd_band, t->grad_band})
        if (band < 1 || band > TBX_MAX_BAND)
            return fail(TBX_ERR_ARG, "band %d outside 1..%d", band, TBX_MAX_BAND);
    return TBX_OK;
}

static int build_mols(const HostSpecies &hs, int nmol, const int *mol_nat, const int *id,
                      HostMols &hm)
{
    if (nmol <= 0 || !mol_nat || !id) return fail(TBX_ERR_ARG, "empty batch");
    hm.nmol = nmol;
    hm.trans.assign(3, 0.0);
    hm.mol_tr.assign(nmol, make_int2(0, 1));
    hm.mol_atom.assign(nmol + 1, 0);
    for (int m = 0; m < nmol; ++m) {
        if (mol_nat[m] <= 0) return fail(TBX_ERR_ARG, "molecule %d has %d atoms", m, mol_nat[m]);
        hm.mol_atom[m + 1] = hm.mol_atom[m] + mol_nat[m];
    }
    const int nat = hm.mol_atom[nmol];
    hm.nat = nat;
    hm.at_sp.assign(id, id + nat);
    hm.at_mol.resize(nat);
    hm.at_sh.resize(nat);
    hm.at_ao.resize(nat);
    hm.at_cart.resize(nat);
    hm.mol_nao.assign(nmol, 0);
    hm.mol_cart.assign(nmol + 1, 0);
    hm.mol_ao.assign(nmol + 1, 0);
    hm.mol_dense.assign(nmol + 1, 0);
    hm.sh_off.assign(nmol + 1, 0);
    int sh = 0, aog = 0, cg = 0;
    for (int m = 0; m < nmol; ++m) {
        int ao = 0, c = 0;
        hm.sh_off[m] = sh;
        for (int a = hm.mol_atom[m]; a < hm.mol_atom[m + 1]; ++a) {
            const int s = hm.at_sp[a];
            if (s < 0 || s >= hs.nsp) return fail(TBX_ERR_ARG, "atom %d: species %d out of range", a, s);
            hm.at_mol[a] = m;
            hm.at_sh[a] = sh;
            hm.at_ao[a] = ao;
            hm.at_cart[a] = c;
            for (int q = 0; q < hs.nsh[s]; ++q)
                hm.sh_aog.push_back(hm.mol_ao[m] + ao + hs.sh_ao[q + hs.msh * s]);
            sh += hs.nsh[s];
            ao += hs.nao[s];
            c += hs.ncart[s];
            for (int k = 0; k < hs.ncart[s]; ++k) hm.cart_atom.push_back(a);
            for (int k = 0; k < hs.nao[s]; ++k) hm.ao_atom.push_back(a);
        }
        hm.mol_nao[m] = ao;
        hm.mol_ao[m + 1] = hm.mol_ao[m] + ao;
        hm.mol_cart[m + 1] = hm.mol_cart[m] + c;
        hm.mol_dense[m + 1] = hm.mol_dense[m] + (long long)ao * ao;
        aog += ao;
        cg += c;
    }
    hm.sh_off[nmol] = sh;
    hm.sh_aog.push_back(aog);
    for (int m = 0; m < nmol; ++m) {  // 32 x 32 tile pairs (bi <= bj) of every dense block
        const int t = (hm.mol_nao[m] + 31) / 32;
        for (int bj = 0; bj < t; ++bj)
            for (int bi = 0; bi <= bj; ++bi) hm.tiles.insert(hm.tiles.end(), {m, bi, bj, 0});
        const int td = (hm.mol_nao[m] + 63) / 64;
        for (int bj = 0; bj < td; ++bj)
            for (int bi = 0; bi <= bj; ++bi) hm.dtiles.insert(hm.dtiles.end(), {m, bi, bj, 0});
    }
    hm.nsh = sh;
    hm.nao = aog;
    hm.ncart = cg;
    hm.ndense = hm.mol_dense[nmol];
    return TBX_OK;
}

// Host staging of many small tables for one host-to-device copy: every
// table starts on a 256-byte boundary of the arena.
struct Arena {
    std::vector<char> host;
    template <class T> size_t add(const std::vector<T> &v)
    {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + sizeof(T) * std::max<size_t>(v.size(), 1), 0);
        if (!v.empty()) std::memcpy(host.data() + off, v.data(), sizeof(T) * v.size());
        return off;
    }
};

// Device copies of the tables: one buffer for the species tables, one for
// the molecule and work-unit tables.  The host arenas stay alive with them.
struct DevTables {
    DevBuf species, mols;
    Arena sp_host, mol_host;
    TbxSpecies sp{};
    TbxMols mt{};

    void release()
    {
        species.release();
        mols.release();
    }
};

static hipError_t upload_arena(DevBuf &d, const Arena &a, hipStream_t s)
{
    hipError_t e = d.reserve(a.host.size());
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(d.p, a.host.data(), a.host.size(), hipMemcpyHostToDevice, s);
}

static hipError_t upload_species(DevTables &d, const HostSpecies &hs, double cutoff, hipStream_t s)
{
    Arena &a = d.sp_host;
    a.host.clear();
    const size_t o_nsh = a.add(hs.nsh), o_ang = a.add(hs.ang), o_nprim = a.add(hs.nprim),
                 o_alpha = a.add(hs.alpha), o_coeff = a.add(hs.coeff), o_shpoly = a.add(hs.shpoly),
                 o_rad = a.add(hs.rad), o_hscale = a.add(hs.hscale), o_ncart = a.add(hs.ncart),
                 o_nao = a.add(hs.nao), o_cart_sh = a.add(hs.cart_sh),
                 o_cart_exp = a.add(hs.cart_exp), o_ao_sh = a.add(hs.ao_sh), o_ao_m = a.add(hs.ao_m),
                 o_sh_cart = a.add(hs.sh_cart), o_sh_ao = a.add(hs.sh_ao), o_prim = a.add(hs.prim),
                 o_prim_off = a.add(hs.prim_off), o_pp = a.add(hs.pp), o_pp_off = a.add(hs.pp_off),
                 o_pp_mu = a.add(hs.pp_mu), o_ao_info = a.add(hs.ao_info),
                 o_ao_desc = a.add(hs.ao_desc), o_onsite = a.add(hs.onsite),
                 o_kdiat = hs.diat ? a.add(hs.kdiat) : 0;
    const hipError_t e = upload_arena(d.species, a, s);
    if (e != hipSuccess) return e;
    const char *base = static_cast<const char *>(d.species.p);
    auto at = [base](size_t off) { return base + off; };
    TbxSpecies &sp = d.sp;
    sp.nsp = hs.nsp;
    sp.msh = hs.msh;
    sp.nsh = reinterpret_cast<const int *>(at(o_nsh));
    sp.ang = reinterpret_cast<const int *>(at(o_ang));
    sp.nprim = reinterpret_cast<const int *>(at(o_nprim));
    sp.alpha = reinterpret_cast<const double *>(at(o_alpha));
    sp.coeff = reinterpret_cast<const double *>(at(o_coeff));
    sp.shpoly = reinterpret_cast<const double *>(at(o_shpoly));
    sp.rad = reinterpret_cast<const double *>(at(o_rad));
    sp.hscale = reinterpret_cast<const double *>(at(o_hscale));
    sp.ncart = reinterpret_cast<const int *>(at(o_ncart));
    sp.nao = reinterpret_cast<const int *>(at(o_nao));
    sp.cart_sh = reinterpret_cast<const int *>(at(o_cart_sh));
    sp.cart_exp = reinterpret_cast<const int *>(at(o_cart_exp));
    sp.ao_sh = reinterpret_cast<const int *>(at(o_ao_sh));
    sp.ao_m = reinterpret_cast<const int *>(at(o_ao_m));
    sp.sh_cart = reinterpret_cast<const int *>(at(o_sh_cart));
    sp.sh_ao = reinterpret_cast<const int *>(at(o_sh_ao));
    sp.prim = reinterpret_cast<const double *>(at(o_prim));
    sp.prim_off = reinterpret_cast<const int *>(at(o_prim_off));
    sp.nprim_tot = (int)(hs.prim.size() / 2);
    sp.nslot = hs.nsp * hs.msh;
    sp.pp_off = reinterpret_cast<const int *>(at(o_pp_off));
    sp.pp = reinterpret_cast<const double *>(at(o_pp));
    sp.pp_mu = reinterpret_cast<const double *>(at(o_pp_mu));
    sp.onsite = reinterpret_cast<const double *>(at(o_onsite));
    sp.ao_info = reinterpret_cast<const int *>(at(o_ao_info));
    sp.ao_desc = reinterpret_cast<const int *>(at(o_ao_desc));
    sp.kdiat = hs.diat ? reinterpret_cast<const double *>(at(o_kdiat)) : nullptr;
    sp.intcut = hs.intcut;
    sp.cutoff2 = cutoff * cutoff;
    return hipSuccess;
}

static hipError_t upload_mols(DevTables &d, const HostMols &hm, hipStream_t s)
{
    Arena &a = d.mol_host;
    a.host.clear();
    const std::vector<double> &trans = hm.trans;
    const size_t o_mol_tr = a.add(hm.mol_tr);
    const size_t o_trans = a.add(trans), o_mol_atom = a.add(hm.mol_atom),
                 o_mol_nao = a.add(hm.mol_nao), o_mol_cart = a.add(hm.mol_cart),
                 o_mol_dense = a.add(hm.mol_dense), o_at_sp = a.add(hm.at_sp),
                 o_at_mol = a.add(hm.at_mol), o_at_sh = a.add(hm.at_sh), o_at_ao = a.add(hm.at_ao),
                 o_at_cart = a.add(hm.at_cart), o_cart_atom = a.add(hm.cart_atom),
                 o_ao_atom = a.add(hm.ao_atom), o_mol_ao = a.add(hm.mol_ao),
                 o_uhead = a.add(hm.uhead), o_uplan = a.add(hm.uplan), o_fhead = a.add(hm.fhead),
                 o_fplan = a.add(hm.fplan), o_sh_aog = a.add(hm.sh_aog), o_tiles = a.add(hm.tiles), o_dtiles = a.add(hm.dtiles),
                 o_hint = a.add(hm.hint_xyz);
    const hipError_t e = upload_arena(d.mols, a, s);
    if (e != hipSuccess) return e;
    const char *base = static_cast<const char *>(d.mols.p);
    auto at = [base](size_t off) { return base + off; };
    TbxMols &mt = d.mt;
    mt.nmol = hm.nmol;
    mt.nat = hm.nat;
    mt.ntrans = (int)(trans.size() / 3);
    mt.zero_image = 1;
    for (const int2 &tr : hm.mol_tr)
        if (tr.y != 1 || trans[3 * tr.x] != 0.0 || trans[3 * tr.x + 1] != 0.0 ||
            trans[3 * tr.x + 2] != 0.0)
            mt.zero_image = 0;
    mt.trans = reinterpret_cast<const double *>(at(o_trans));
    mt.mol_tr = reinterpret_cast<const int2 *>(at(o_mol_tr));
    mt.mol_atom = reinterpret_cast<const int *>(at(o_mol_atom));
    mt.mol_nao = reinterpret_cast<const int *>(at(o_mol_nao));
    mt.mol_cart = reinterpret_cast<const int *>(at(o_mol_cart));
    mt.mol_dense = reinterpret_cast<const long long *>(at(o_mol_dense));
    mt.at_sp = reinterpret_cast<const int *>(at(o_at_sp));
    mt.at_mol = reinterpret_cast<const int *>(at(o_at_mol));
    mt.at_sh = reinterpret_cast<const int *>(at(o_at_sh));
    mt.at_ao = reinterpret_cast<const int *>(at(o_at_ao));
    mt.at_cart = reinterpret_cast<const int *>(at(o_at_cart));
    mt.cart_atom = reinterpret_cast<const int *>(at(o_cart_atom));
    mt.ao_atom = reinterpret_cast<const int *>(at(o_ao_atom));
    mt.mol_ao = reinterpret_cast<const int *>(at(o_mol_ao));
    mt.nunit = (int)hm.uhead.size();
    mt.unit = reinterpret_cast<const TbxUnitHead *>(at(o_uhead));
    mt.plan = reinterpret_cast<const int *>(at(o_uplan));
    mt.nfunit = (int)hm.fhead.size();
    mt.funit = reinterpret_cast<const TbxUnitHead *>(at(o_fhead));
    mt.fplan = reinterpret_cast<const int *>(at(o_fplan));
    mt.nsh_tot = hm.nsh;
    mt.nao_tot = hm.nao;
    mt.sh_aog = reinterpret_cast<const int *>(at(o_sh_aog));
    mt.ntile = (int)(hm.tiles.size() / 4);
    mt.tiles = reinterpret_cast<const int *>(at(o_tiles));
    mt.ndtile = (int)(hm.dtiles.size() / 4);
    mt.dtiles = reinterpret_cast<const int *>(at(o_dtiles));
    // the far gradient units may be skipped while no atom moved more than
    // (a little less than) half the skin away from the hint
    const bool hinted = hm.hint_xyz.size() == 3 * (size_t)hm.nat && hm.nat > 0;
    mt.nunit_near = hinted ? hm.nunit_near : mt.nunit;
    mt.nfunit_far = hinted ? hm.nfunit_far : 0;
    mt.hint = hinted ? reinterpret_cast<const double *>(at(o_hint)) : nullptr;
    mt.hint_r2 = 0.25 * 0.999 * hm.skin * hm.skin;
    return hipSuccess;
}

// ---------------------------------------------------------------------------

struct tbx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    Budget budget;  // device memory of the context and its batches
    DevTables tab;
    DevBuf xyz, se, dsedcn, S, D, Q, H, P, W, vao, vdp, vqp, grad, dse, sigma, dEdcn, nnl, flag;
    std::vector<double> host_tmp;
    // species tables of the last call (host copy, and the raw inputs they were
    // built from): calls with the same basis, H0 parameters and cutoff reuse
    // them and their device copy
    HostSpecies hs;
    std::vector<char> hs_raw;
    bool hs_valid = false;
    // molecule tables and work-unit plans of the last call (they depend on the
    // species, the atom order and the translations, not on the coordinates):
    // a call with the same species tables, ids and translations reuses them
    // and their device copy
    HostMols hm;
    std::vector<char> hm_raw;
    bool hm_valid = false;
};

// The raw bytes of a basis + H0 parameter set + cutoff (empty if a table is
// missing), to recognise a repeated parametrisation exactly.
static std::vector<char> species_bytes(const tbx_basis *bas, const tbx_h0par *h0, double cutoff)
{
    std::vector<char> r;
    if (!bas || !h0 || bas->nid <= 0 || bas->msh <= 0 || bas->msh > TBX_MAXSH || !bas->nsh_id ||
        !bas->ang || !bas->nprim || !bas->alpha || !bas->coeff || !h0->hscale || !h0->shpoly ||
        !h0->rad)
        return r;
    const size_t nsp = bas->nid, msh = bas->msh;
    auto put = [&r](const void *p, size_t n) {
        const char *c = static_cast<const char *>(p);
        r.insert(r.end(), c, c + n);
    };
    put(&bas->nid, sizeof(int));
    put(&bas->msh, sizeof(int));
    put(&bas->intcut, sizeof(double));
    put(&cutoff, sizeof(double));
    put(bas->nsh_id, sizeof(int) * nsp);
    put(bas->ang, sizeof(int) * nsp * msh);
    put(bas->nprim, sizeof(int) * nsp * msh);
    put(bas->alpha, sizeof(double) * TBX_MAXG * nsp * msh);
    put(bas->coeff, sizeof(double) * TBX_MAXG * nsp * msh);
    put(h0->hscale, sizeof(double) * msh * msh * nsp * nsp);
    put(h0->shpoly, sizeof(double) * msh * nsp);
    put(h0->rad, sizeof(double) * nsp);
    const int diat = (h0->ksig && h0->kpi && h0->kdel) ? 1 : 0;
    if (!diat && (h0->ksig || h0->kpi || h0->kdel)) return std::vector<char>();
    put(&diat, sizeof(int));
    if (diat) {
        put(h0->ksig, sizeof(double) * nsp * nsp);
        put(h0->kpi, sizeof(double) * nsp * nsp);
        put(h0->kdel, sizeof(double) * nsp * nsp);
    }
    return r;
}

struct tbx_batch {
    tbx_context *ctx = nullptr;
    HostSpecies hs;
    HostMols hm;
    tbx_tiles tiles = kDefaultTiles;
    double cutoff = 0.0;  // neighbour cutoff (bohr)
    DevTables tab;
    DevBuf flag, dse_tmp;
    hipEvent_t ev[16] = {};  // pairs (start, end): forward, -, gradient, adjoint, 4 SCF kernels
    bool timed = false;
    // generalized eigensolver (rocSOLVER): handle, copy of S, workspace, info
    rocblas_handle solver = nullptr;
    DevBuf sy_b, sy_e, sy_info;
    // coordination number: covalent radii per species (device) and constants
    DevBuf rcov;
    bool cn_set = false;
    double cn_ka = 0.0, cn_kb = 0.0, cn_rshift = 0.0, cn_cutoff = 0.0;
};

extern "C" {

int tbx_version(void) { return 100; }

const char *tbx_last_error(void) { return g_error.c_str(); }

tbx_context *tbx_context_new(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        fail(TBX_ERR_DEVICE, "no HIP device %d (count %d)", device, n);
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        fail(TBX_ERR_DEVICE, "hipSetDevice(%d) failed", device);
        return nullptr;
    }
    if (scratch_init(device) != TBX_OK) return nullptr;
    tbx_context *ctx = new tbx_context;
    ctx->device = device;
    for (DevBuf *b : {&ctx->tab.species, &ctx->tab.mols, &ctx->xyz, &ctx->se, &ctx->dsedcn, &ctx->S,
                      &ctx->D, &ctx->Q, &ctx->H, &ctx->P, &ctx->W, &ctx->vao, &ctx->vdp, &ctx->vqp,
                      &ctx->grad, &ctx->dse, &ctx->sigma, &ctx->dEdcn, &ctx->nnl, &ctx->flag})
        b->budget = &ctx->budget;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        fail(TBX_ERR_DEVICE, "stream creation failed");
        delete ctx;
        return nullptr;
    }
    return ctx;
}

void tbx_context_delete(tbx_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->tab.release();
    DevBuf *all[] = {&ctx->xyz, &ctx->se, &ctx->dsedcn, &ctx->S, &ctx->D, &ctx->Q, &ctx->H,
                     &ctx->P, &ctx->W, &ctx->vao, &ctx->vdp, &ctx->vqp, &ctx->grad, &ctx->dse,
                     &ctx->sigma, &ctx->dEdcn, &ctx->nnl, &ctx->flag};
    for (DevBuf *b : all) b->release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int tbx_context_set_memory_limit(tbx_context *ctx, long long bytes)
{
    if (!ctx) return fail(TBX_ERR_ARG, "null context");
    ctx->budget.limit = bytes < 0 ? SIZE_MAX : (size_t)bytes;
    return TBX_OK;
}

long long tbx_context_memory_used(const tbx_context *ctx)
{
    return ctx ? (long long)ctx->budget.used : -1;
}

int tbx_get_selfenergy(int nat, const int *id, const int *ish_at, const int *nsh_id, int msh,
                       const double *selfenergy_sp, const double *kcn, const double *cn,
                       double *selfenergy, double *dsedcn)
{
    if (nat < 0 || !id || !ish_at || !nsh_id || !selfenergy_sp || !selfenergy || (cn && !kcn))
        return fail(TBX_ERR_ARG, "tbx_get_selfenergy: invalid argument");
    for (int iat = 0; iat < nat; ++iat) {
        const int izp = id[iat], ii = ish_at[iat];
        for (int ish = 0; ish < nsh_id[izp]; ++ish) {
            double se = selfenergy_sp[ish + msh * izp];
            if (cn) se -= kcn[ish + msh * izp] * cn[iat];
            selfenergy[ii + ish] = se;
            if (dsedcn) dsedcn[ii + ish] = cn ? -kcn[ish + msh * izp] : 0.0;
        }
    }
    return TBX_OK;
}

// get_lattice_points_cutoff (cutoff.f90:97-121): the repetitions along each
// lattice vector follow from the distances between lattice planes
// (get_translations, :127-160) and the points are listed in the order of
// get_lattice_points_rep_3d with the origin first (:39-94).
int tbx_get_lattice_points(const int *periodic, const double *lattice, double rthr, int *ntrans,
                           double *trans)
{
    if (!periodic || !lattice || !ntrans || (*ntrans > 0 && !trans))
        return fail(TBX_ERR_ARG, "tbx_get_lattice_points: null argument");
    const int cap = *ntrans;
    if (!periodic[0] && !periodic[1] && !periodic[2]) {
        *ntrans = 1;
        if (cap < 1) return fail(TBX_ERR_ARG, "room for 1 translation needed");
        trans[0] = trans[1] = trans[2] = 0.0;
        return TBX_OK;
    }
    if (!(rthr >= 0.0) || !std::isfinite(rthr)) return fail(TBX_ERR_ARG, "cutoff %g", rthr);
    const double *a[3] = {lattice, lattice + 3, lattice + 6};
    int rep[3];
    for (int k = 0; k < 3; ++k) {
        const double *u = a[(k + 1) % 3], *v = a[(k + 2) % 3];
        const double n[3] = {u[1] * v[2] - v[1] * u[2], u[2] * v[0] - v[2] * u[0],
                             u[0] * v[1] - v[0] * u[1]};
        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        const double c = (n[0] * a[k][0] + n[1] * a[k][1] + n[2] * a[k][2]) / nn;
        const double r = std::ceil(std::fabs(rthr / c));
        if (!std::isfinite(r) || r > 1000.0)
            return fail(TBX_ERR_ARG, "degenerate lattice or cutoff: %g repetitions", r);
        rep[k] = (int)r;
    }
    const long long need = (2LL * rep[0] + 1) * (2LL * rep[1] + 1) * (2LL * rep[2] + 1);
    if (need > INT_MAX) return fail(TBX_ERR_ARG, "too many lattice points");
    *ntrans = (int)need;
    if (cap < need) return fail(TBX_ERR_ARG, "room for %lld translations needed", need);
    int itr = 0;
    for (int ix = 0; ix <= rep[0]; ++ix)
        for (int iy = 0; iy <= rep[1]; ++iy)
            for (int iz = 0; iz <= rep[2]; ++iz)
                for (int jx = 1; jx >= (ix > 0 ? -1 : 1); jx -= 2)
                    for (int jy = 1; jy >= (iy > 0 ? -1 : 1); jy -= 2)
                        for (int jz = 1; jz >= (iz > 0 ? -1 : 1); jz -= 2) {
                            for (int c = 0; c < 3; ++c)
                                trans[3 * itr + c] = a[0][c] * ix * jx + a[1][c] * iy * jy +
                                                     a[2][c] * iz * jz;
                            ++itr;
                        }
    return TBX_OK;
}

} // extern "C"

// The raw bytes of a structure's species ids and translations.
static std::vector<char> mols_bytes(const tbx_geom *geom)
{
    std::vector<char> r;
    auto put = [&r](const void *p, size_t n) {
        const char *c = static_cast<const char *>(p);
        r.insert(r.end(), c, c + n);
    };
    put(&geom->nat, sizeof(int));
    put(&geom->ntrans, sizeof(int));
    put(geom->id, sizeof(int) * (size_t)geom->nat);
    put(geom->trans, sizeof(double) * 3 * (size_t)geom->ntrans);
    return r;
}

// Shared set-up of the single-molecule host path: the context's species
// tables (ctx->hs) and molecule tables with work-unit plans (ctx->hm), built
// and uploaded only when the parametrisation, the species of the atoms or the
// translations change.
static int prepare_single(tbx_context *ctx, const tbx_geom *geom, const tbx_list *list,
                          const tbx_basis *bas, const tbx_h0par *h0, const int *ish_at,
                          const int *iao_sh, int nao)
{
    if (!ctx) return fail(TBX_ERR_ARG, "null context");
    if (!geom || geom->nat <= 0 || !geom->id || !geom->xyz)
        return fail(TBX_ERR_ARG, "invalid geometry");
    if (geom->ntrans < 1 || !geom->trans) return fail(TBX_ERR_ARG, "need at least one translation");
    if (!(geom->cutoff > 0.0) || !std::isfinite(geom->cutoff))
        return fail(TBX_ERR_ARG, "invalid cutoff %g", geom->cutoff);
    for (int k = 0; k < 3 * geom->nat; ++k)
        if (!std::isfinite(geom->xyz[k])) return fail(TBX_ERR_ARG, "non-finite coordinate %d", k);
    // species tables: rebuilt and uploaded only when the parametrisation changes
    std::vector<char> raw = species_bytes(bas, h0, geom->cutoff);
    const bool reuse = ctx->hs_valid && !raw.empty() && raw == ctx->hs_raw;
    int st = TBX_OK;
    HostSpecies hs_new;
    if (!reuse) {
        ctx->hs_valid = false;
        ctx->hm_valid = false;
        if ((st = build_species(bas, h0, hs_new))) return st;
    }
    const HostSpecies &hs = reuse ? ctx->hs : hs_new;
    std::vector<char> mraw = mols_bytes(geom);
    const bool reuse_mols = reuse && ctx->hm_valid && mraw == ctx->hm_raw;
    HostMols hm_new;
    if (!reuse_mols) {
        ctx->hm_valid = false;
        if ((st = build_mols(hs, 1, &geom->nat, geom->id, hm_new))) return st;
    }
    const HostMols &hm = reuse_mols ? ctx->hm : hm_new;
    if (hm.nao != nao) return fail(TBX_ERR_ARG, "nao=%d but the basis gives %d", nao, hm.nao);
    if (ish_at && iao_sh) {
        for (int a = 0; a < geom->nat; ++a) {
            if (ish_at[a] != hm.at_sh[a])
                return fail(TBX_ERR_ARG, "ish_at[%d]=%d, expected %d", a, ish_at[a], hm.at_sh[a]);
            const int s = hm.at_sp[a];
            for (int k = 0; k < hs.nsh[s]; ++k)
                if (iao_sh[ish_at[a] + k] != hm.at_ao[a] + hs.sh_ao[k + hs.msh * s])
                    return fail(TBX_ERR_ARG, "iao_sh does not follow the basis order");
        }
    }
    if (list && (!list->nnl || !list->inl)) return fail(TBX_ERR_ARG, "incomplete neighbour list");
    hipStream_t s = ctx->stream;
    HIP_TRY(hipSetDevice(ctx->device));
    if (!reuse) {
        // no launch of this context may still read the tables replaced here
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(upload_species(ctx->tab, hs_new, geom->cutoff, s));
        ctx->hs = std::move(hs_new);
        ctx->hs_raw = std::move(raw);
        ctx->hs_valid = true;
    }
    if (!reuse_mols) {
        hm_new.trans.assign(geom->trans, geom->trans + 3 * geom->ntrans);
        hm_new.mol_tr.assign(1, make_int2(0, geom->ntrans));
        if ((st = build_units(ctx->hs, hm_new, kDefaultTiles))) return st;
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(upload_mols(ctx->tab, hm_new, s));
        ctx->hm = std::move(hm_new);
        ctx->hm_raw = std::move(mraw);
        ctx->hm_valid = true;
    }
    HIP_TRY(ctx->xyz.reserve(sizeof(double) * 3 * geom->nat));
    HIP_TRY(hipMemcpyAsync(ctx->xyz.p, geom->xyz, sizeof(double) * 3 * geom->nat,
                           hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->flag.reserve(sizeof(int)));
    HIP_TRY(hipMemsetAsync(ctx->flag.p, 0, sizeof(int), s));
    if (list) {
        HIP_TRY(ctx->nnl.reserve(sizeof(int) * geom->nat));
        HIP_TRY(hipMemcpyAsync(ctx->nnl.p, list->nnl, sizeof(int) * geom->nat,
                               hipMemcpyHostToDevice, s));
    }
    return TBX_OK;
}

static int check_flag(tbx_context *ctx)
{
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, ctx->flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (flag)
        return fail(TBX_ERR_ARG,
                    "neighbour list does not match the cutoff rule of new_adjacency_list");
    return TBX_OK;
}

extern "C" int tbx_get_hamiltonian(tbx_context *ctx, const tbx_geom *geom, const tbx_list *list,
                                   const tbx_basis *bas, const tbx_h0par *h0, const int *ish_at,
                                   const int *iao_sh, int nao, const double *selfenergy,
                                   double *overlap, double *dpint, double *qpint,
                                   double *hamiltonian)
{
    int st = prepare_single(ctx, geom, list, bas, h0, ish_at, iao_sh, nao);
    if (st) return st;
    const HostMols &hm = ctx->hm;
    if (!selfenergy || !overlap || !dpint || !qpint || !hamiltonian)
        return fail(TBX_ERR_ARG, "null output array");
    hipStream_t s = ctx->stream;
    const size_t n2 = (size_t)nao * nao;
    HIP_TRY(ctx->se.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(hipMemcpyAsync(ctx->se.p, selfenergy, sizeof(double) * hm.nsh, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->S.reserve(sizeof(double) * n2));
    HIP_TRY(ctx->D.reserve(sizeof(double) * 3 * n2));
    HIP_TRY(ctx->Q.reserve(sizeof(double) * 6 * n2));
    HIP_TRY(ctx->H.reserve(sizeof(double) * n2));
    HIP_TRY(tbx_launch_forward(&ctx->tab.sp, &ctx->tab.mt, ctx->xyz.as<double>(),
                               ctx->se.as<double>(), ctx->S.as<double>(), ctx->D.as<double>(),
                               ctx->Q.as<double>(), ctx->H.as<double>(), hm.ndense,
                               list ? ctx->nnl.as<int>() : nullptr, ctx->flag.as<int>(), 0, s));
    HIP_TRY(hipMemcpyAsync(overlap, ctx->S.p, sizeof(double) * n2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(dpint, ctx->D.p, sizeof(double) * 3 * n2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(qpint, ctx->Q.p, sizeof(double) * 6 * n2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hamiltonian, ctx->H.p, sizeof(double) * n2, hipMemcpyDeviceToHost, s));
    return check_flag(ctx);
}

// get_multipole_integrals / get_dipole_integrals / get_overlap: the forward
// kernel over the pairs within geom->cutoff with neutral H0 parameters (the
// H0 it writes to device scratch is dropped).  dpint / qpint may be NULL.
extern "C" int tbx_get_multipole_integrals(tbx_context *ctx, const tbx_geom *geom,
                                           const tbx_basis *bas, const int *ish_at,
                                           const int *iao_sh, int nao, double *overlap,
                                           double *dpint, double *qpint)
{
    if (!bas || bas->nid <= 0 || bas->msh <= 0 || bas->msh > TBX_MAXSH)
        return fail(TBX_ERR_BASIS, "invalid basis");
    if (!overlap) return fail(TBX_ERR_ARG, "null output array");
    const int nid = bas->nid, msh = bas->msh;
    std::vector<double> hscale((size_t)msh * msh * nid * nid, 0.0), shpoly((size_t)msh * nid, 0.0),
        rad(nid, 1.0);
    const tbx_h0par h0 = {hscale.data(), shpoly.data(), rad.data()};
    int st = prepare_single(ctx, geom, nullptr, bas, &h0, ish_at, iao_sh, nao);
    if (st) return st;
    const HostMols &hm = ctx->hm;
    hipStream_t s = ctx->stream;
    const size_t n2 = (size_t)nao * nao;
    const std::vector<double> se(hm.nsh, 0.0);
    HIP_TRY(ctx->se.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(hipMemcpyAsync(ctx->se.p, se.data(), sizeof(double) * hm.nsh, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->S.reserve(sizeof(double) * n2));
    HIP_TRY(ctx->D.reserve(sizeof(double) * 3 * n2));
    HIP_TRY(ctx->Q.reserve(sizeof(double) * 6 * n2));
    HIP_TRY(ctx->H.reserve(sizeof(double) * n2));
    HIP_TRY(tbx_launch_forward(&ctx->tab.sp, &ctx->tab.mt, ctx->xyz.as<double>(),
                               ctx->se.as<double>(), ctx->S.as<double>(), ctx->D.as<double>(),
                               ctx->Q.as<double>(), ctx->H.as<double>(), hm.ndense, nullptr,
                               ctx->flag.as<int>(), 0, s));
    HIP_TRY(hipMemcpyAsync(overlap, ctx->S.p, sizeof(double) * n2, hipMemcpyDeviceToHost, s));
    if (dpint)
        HIP_TRY(hipMemcpyAsync(dpint, ctx->D.p, sizeof(double) * 3 * n2, hipMemcpyDeviceToHost, s));
    if (qpint)
        HIP_TRY(hipMemcpyAsync(qpint, ctx->Q.p, sizeof(double) * 6 * n2, hipMemcpyDeviceToHost, s));
    return check_flag(ctx);
}

extern "C" int tbx_get_dipole_integrals(tbx_context *ctx, const tbx_geom *geom,
                                        const tbx_basis *bas, const int *ish_at, const int *iao_sh,
                                        int nao, double *overlap, double *dpint)
{
    if (!dpint) return fail(TBX_ERR_ARG, "null output array");
    return tbx_get_multipole_integrals(ctx, geom, bas, ish_at, iao_sh, nao, overlap, dpint,
                                       nullptr);
}

extern "C" int tbx_get_overlap(tbx_context *ctx, const tbx_geom *geom, const tbx_basis *bas,
                               const int *ish_at, const int *iao_sh, int nao, double *overlap)
{
    return tbx_get_multipole_integrals(ctx, geom, bas, ish_at, iao_sh, nao, overlap, nullptr,
                                       nullptr);
}

extern "C" int tbx_get_hamiltonian_gradient(tbx_context *ctx, const tbx_geom *geom,
                                            const tbx_list *list, const tbx_basis *bas,
                                            const tbx_h0par *h0, const int *ish_at,
                                            const int *iao_sh, int nao, const double *selfenergy,
                                            const double *dsedcn, int nspin, const double *vao,
                                            const double *vdp, const double *vqp,
                                            const double *pmat, const double *xmat,
                                            double *dEdcn, double *gradient, double *sigma)
{
    int st = prepare_single(ctx, geom, list, bas, h0, ish_at, iao_sh, nao);
    if (st) return st;
    if (ctx->hs.diat) return fail(TBX_ERR_ARG, kNoDiatGradient);
    const HostMols &hm = ctx->hm;
    if (nspin < 1 || nspin > 2) return fail(TBX_ERR_ARG, "nspin=%d", nspin);
    if (!selfenergy || !dsedcn || !vao || !vdp || !vqp || !pmat || !xmat || !dEdcn || !gradient ||
        !sigma)
        return fail(TBX_ERR_ARG, "null argument");
    hipStream_t s = ctx->stream;
    const int nat = geom->nat;
    const size_t n2 = (size_t)nao * nao;
    HIP_TRY(ctx->se.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(hipMemcpyAsync(ctx->se.p, selfenergy, sizeof(double) * hm.nsh, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->dsedcn.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(hipMemcpyAsync(ctx->dsedcn.p, dsedcn, sizeof(double) * hm.nsh, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->P.reserve(sizeof(double) * n2 * nspin));
    HIP_TRY(hipMemcpyAsync(ctx->P.p, pmat, sizeof(double) * n2 * nspin, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->W.reserve(sizeof(double) * n2));
    HIP_TRY(hipMemcpyAsync(ctx->W.p, xmat, sizeof(double) * n2, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->vao.reserve(sizeof(double) * nao * nspin));
    HIP_TRY(hipMemcpyAsync(ctx->vao.p, vao, sizeof(double) * nao * nspin, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->vdp.reserve(sizeof(double) * 3 * nat));
    HIP_TRY(hipMemcpyAsync(ctx->vdp.p, vdp, sizeof(double) * 3 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->vqp.reserve(sizeof(double) * 6 * nat));
    HIP_TRY(hipMemcpyAsync(ctx->vqp.p, vqp, sizeof(double) * 6 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->grad.reserve(sizeof(double) * 3 * nat));
    HIP_TRY(ctx->dse.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(ctx->sigma.reserve(sizeof(double) * 9));
    HIP_TRY(ctx->dEdcn.reserve(sizeof(double) * nat));
    HIP_TRY(tbx_launch_fill(ctx->grad.as<double>(), 3LL * nat, 0.0, s));
    HIP_TRY(tbx_launch_fill(ctx->dse.as<double>(), hm.nsh, 0.0, s));
    HIP_TRY(tbx_launch_fill(ctx->sigma.as<double>(), 9, 0.0, s));
    HIP_TRY(tbx_launch_gradient(0, &ctx->tab.sp, &ctx->tab.mt, ctx->xyz.as<double>(),
                                ctx->se.as<double>(), ctx->P.as<double>(), ctx->W.as<double>(), ctx->vao.as<double>(),
                                ctx->vdp.as<double>(), ctx->vqp.as<double>(), nspin, (long long)n2,
                                nao, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                ctx->grad.as<double>(),
                                ctx->dse.as<double>(), ctx->sigma.as<double>(),
                                list ? ctx->nnl.as<int>() : nullptr, ctx->flag.as<int>(), 0, s));
    HIP_TRY(tbx_launch_dedcn(&ctx->tab.sp, &ctx->tab.mt, ctx->dse.as<double>(),
                             ctx->dsedcn.as<double>(), ctx->P.as<double>(),
                             ctx->dEdcn.as<double>(), s));
    std::vector<double> g(3 * nat), c(nat), sg(9);
    HIP_TRY(hipMemcpyAsync(g.data(), ctx->grad.p, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(c.data(), ctx->dEdcn.p, sizeof(double) * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(sg.data(), ctx->sigma.p, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
    st = check_flag(ctx);  // synchronises the stream
    if (st) return st;
    for (int k = 0; k < 3 * nat; ++k) gradient[k] += g[k];
    for (int k = 0; k < nat; ++k) dEdcn[k] += c[k];
    for (int k = 0; k < 9; ++k) sigma[k] += sg[k];
    return TBX_OK;
}

// Adjoint of get_hamiltonian for one structure: the adjoint kernel of the
// batched path on a one-molecule batch, host arrays in and out.
extern "C" int tbx_get_hamiltonian_adjoint(tbx_context *ctx, const tbx_geom *geom,
                                           const tbx_list *list, const tbx_basis *bas,
                                           const tbx_h0par *h0, const int *ish_at,
                                           const int *iao_sh, int nao, const double *selfenergy,
                                           const double *sbar, const double *hbar,
                                           const double *dbar, const double *qbar,
                                           double *gradient, double *dEdse, double *sigma)
{
    if (!selfenergy || !sbar || !hbar || !dbar || !qbar || !gradient || !dEdse)
        return fail(TBX_ERR_ARG, "null argument");
    int st = prepare_single(ctx, geom, list, bas, h0, ish_at, iao_sh, nao);
    if (st) return st;
    if (ctx->hs.diat) return fail(TBX_ERR_ARG, kNoDiatGradient);
    const HostMols &hm = ctx->hm;
    hipStream_t s = ctx->stream;
    const int nat = geom->nat;
    const size_t n2 = (size_t)nao * nao;
    HIP_TRY(ctx->se.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(hipMemcpyAsync(ctx->se.p, selfenergy, sizeof(double) * hm.nsh, hipMemcpyHostToDevice, s));
    // the adjoints go into the buffers of the forward's outputs
    HIP_TRY(ctx->S.reserve(sizeof(double) * n2));
    HIP_TRY(ctx->H.reserve(sizeof(double) * n2));
    HIP_TRY(ctx->D.reserve(sizeof(double) * 3 * n2));
    HIP_TRY(ctx->Q.reserve(sizeof(double) * 6 * n2));
    HIP_TRY(hipMemcpyAsync(ctx->S.p, sbar, sizeof(double) * n2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->H.p, hbar, sizeof(double) * n2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->D.p, dbar, sizeof(double) * 3 * n2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->Q.p, qbar, sizeof(double) * 6 * n2, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx->grad.reserve(sizeof(double) * 3 * nat));
    HIP_TRY(ctx->dse.reserve(sizeof(double) * hm.nsh));
    HIP_TRY(ctx->sigma.reserve(sizeof(double) * 9));
    HIP_TRY(tbx_launch_fill(ctx->grad.as<double>(), 3LL * nat, 0.0, s));
    HIP_TRY(tbx_launch_fill(ctx->dse.as<double>(), hm.nsh, 0.0, s));
    HIP_TRY(tbx_launch_fill(ctx->sigma.as<double>(), 9, 0.0, s));
    HIP_TRY(tbx_launch_gradient(1, &ctx->tab.sp, &ctx->tab.mt, ctx->xyz.as<double>(),
                                ctx->se.as<double>(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                1, 0, 0, ctx->S.as<double>(), ctx->D.as<double>(),
                                ctx->Q.as<double>(), ctx->H.as<double>(), nullptr, nullptr,
                                ctx->grad.as<double>(),
                                ctx->dse.as<double>(), ctx->sigma.as<double>(),
                                list ? ctx->nnl.as<int>() : nullptr, ctx->flag.as<int>(), 0, s));
    std::vector<double> g(3 * nat), e(hm.nsh), sg(9);
    HIP_TRY(hipMemcpyAsync(g.data(), ctx->grad.p, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(e.data(), ctx->dse.p, sizeof(double) * hm.nsh, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(sg.data(), ctx->sigma.p, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
    st = check_flag(ctx);  // synchronises the stream
    if (st) return st;
    for (int k = 0; k < 3 * nat; ++k) gradient[k] += g[k];
    for (int k = 0; k < hm.nsh; ++k) dEdse[k] += e[k];
    if (sigma)
        for (int k = 0; k < 9; ++k) sigma[k] += sg[k];
    return TBX_OK;
}

// ---------------------------------------------------------------------------
// batched path

// Batch set-up shared by tbx_batch_create and tbx_batch_create_periodic:
// mol_ntrans / trans NULL for molecules (one zero translation each).
static int batch_create(tbx_context *ctx, const tbx_batch_desc *desc, const int *mol_ntrans,
                        const double *trans, tbx_batch **out)
{
    if (!ctx || !desc || !out) return fail(TBX_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(desc->cutoff > 0.0) || !std::isfinite(desc->cutoff))
        return fail(TBX_ERR_ARG, "invalid cutoff %g", desc->cutoff);
    tbx_batch *b = new tbx_batch;
    b->ctx = ctx;
    b->cutoff = desc->cutoff;
    for (DevBuf *d : {&b->tab.species, &b->tab.mols, &b->flag, &b->dse_tmp, &b->rcov, &b->sy_b,
                      &b->sy_e, &b->sy_info})
        d->budget = &ctx->budget;
    int st = build_species(desc->bas, desc->h0, b->hs);
    if (!st) st = build_mols(b->hs, desc->nmol, desc->mol_nat, desc->id, b->hm);
    if (!st && mol_ntrans) {
        long long ntot = 0;
        for (int m = 0; m < desc->nmol && !st; ++m) {
            if (mol_ntrans[m] < 1)
                st = fail(TBX_ERR_ARG, "molecule %d has %d translations (>= 1)", m, mol_ntrans[m]);
            ntot += mol_ntrans[m];
        }
        if (!st && (!trans || ntot > INT_MAX / 3)) st = fail(TBX_ERR_ARG, "invalid translations");
        for (long long k = 0; !st && k < 3 * ntot; ++k)
            if (!std::isfinite(trans[k])) st = fail(TBX_ERR_ARG, "non-finite translation %lld", k);
        if (!st) {
            b->hm.trans.assign(trans, trans + 3 * ntot);
            int first = 0;
            for (int m = 0; m < desc->nmol; ++m) {
                b->hm.mol_tr[m] = make_int2(first, mol_ntrans[m]);
                first += mol_ntrans[m];
            }
        }
    }
    if (st) {
        delete b;
        return st;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = upload_species(b->tab, b->hs, desc->cutoff, ctx->stream);
    if ((st = build_units(b->hs, b->hm, b->tiles))) {
        tbx_batch_delete(b);
        return st;
    }
    if (e == hipSuccess) e = upload_mols(b->tab, b->hm, ctx->stream);
    if (e == hipSuccess) e = b->flag.reserve(sizeof(int));
    // scratch of the gradient / adjoint launches: dE/dse (plain) or sigma (adjoint)
    if (e == hipSuccess)
        e = b->dse_tmp.reserve(sizeof(double) * std::max({1, b->hm.nsh, 9 * b->hm.nmol}));
    if (e == hipSuccess) e = hipMemsetAsync(b->flag.p, 0, sizeof(int), ctx->stream);
    for (int k = 0; k < 16 && e == hipSuccess; ++k) e = hipEventCreate(&b->ev[k]);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    b->tab.sp_host.host = std::vector<char>();  // copies done: free the host staging
    b->tab.mol_host.host = std::vector<char>();
    if (e != hipSuccess) {
        tbx_batch_delete(b);
        return fail(e == hipErrorOutOfMemory ? TBX_ERR_ALLOC : TBX_ERR_DEVICE,
                    "batch upload: %s", hipGetErrorString(e));
    }
    *out = b;
    return TBX_OK;
}

extern "C" int tbx_batch_create(tbx_context *ctx, const tbx_batch_desc *desc, tbx_batch **out)
{
    return batch_create(ctx, desc, nullptr, nullptr, out);
}

extern "C" int tbx_batch_create_periodic(tbx_context *ctx, const tbx_batch_desc *desc,
                                         const int *mol_ntrans, const double *trans,
                                         tbx_batch **out)
{
    if (!mol_ntrans || !trans) return fail(TBX_ERR_ARG, "null translations");
    return batch_create(ctx, desc, mol_ntrans, trans, out);
}

extern "C" int tbx_batch_set_tiles(tbx_batch *b, const tbx_tiles *tiles)
{
    if (!b || !tiles) return fail(TBX_ERR_ARG, "null argument");
    if (int st = check_tiles(b->hs, tiles)) return st;
    if (int st = build_units(b->hs, b->hm, *tiles)) {
        (void)build_units(b->hs, b->hm, b->tiles);  // keep the host copy of the current units
        return st;
    }
    b->tiles = *tiles;
    HIP_TRY(hipSetDevice(b->ctx->device));
    // The batch's kernels run on the callers' streams: wait for every stream
    // of the device, so that no launch still reads the units replaced here.
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(upload_mols(b->tab, b->hm, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->tab.mol_host.host = std::vector<char>();
    return TBX_OK;
}

// Geometry hint of the pair order (see build_units_list); the cutoff is the
// batch's neighbour cutoff.
static int set_hint(HostMols &hm, const double *xyz, double skin, double cutoff)
{
    if (xyz) {  // validate before anything changes
        if (!(skin >= 0.0) || !std::isfinite(skin)) return fail(TBX_ERR_ARG, "skin %g", skin);
        for (int k = 0; k < 3 * hm.nat; ++k)
            if (!std::isfinite(xyz[k]))
                return fail(TBX_ERR_ARG, "coordinate %d is not finite", k);
    }
    hm.hint_xyz.clear();
    hm.near2 = hm.skin = 0.0;
    if (!xyz) return TBX_OK;
    hm.hint_xyz.assign(xyz, xyz + 3 * (size_t)hm.nat);
    hm.near2 = (cutoff + skin) * (cutoff + skin);
    hm.skin = skin;
    return TBX_OK;
}

extern "C" int tbx_batch_order_pairs(tbx_batch *b, const double *xyz, double skin)
{
    if (!b) return fail(TBX_ERR_ARG, "null argument");
    const std::vector<double> old_xyz = b->hm.hint_xyz;
    const double old_near2 = b->hm.near2;
    if (int st = set_hint(b->hm, xyz, skin, b->cutoff)) return st;
    if (int st = build_units(b->hs, b->hm, b->tiles)) {
        b->hm.hint_xyz = old_xyz;  // keep the host copy of the current units
        b->hm.near2 = old_near2;
        (void)build_units(b->hs, b->hm, b->tiles);
        return st;
    }
    HIP_TRY(hipSetDevice(b->ctx->device));
    // The batch's kernels run on the callers' streams: wait for every stream
    // of the device, so that no launch still reads the units replaced here.
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(upload_mols(b->tab, b->hm, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->tab.mol_host.host = std::vector<char>();
    return TBX_OK;
}

extern "C" int tbx_batch_get_tiles(const tbx_batch *b, tbx_tiles *tiles)
{
    if (!b || !tiles) return fail(TBX_ERR_ARG, "null argument");
    *tiles = b->tiles;
    return TBX_OK;
}

// Sections of a staged plan (16-byte words from TbxUnitHead::plan), as the
// kernels' load_unit reads them: pairs, pext, segments, shell pairs, offsets.
struct PlanView {
    const int *pairs, *pext, *sinfo, *spl, *toff, *tsp;
    int npair, nseg, nspair, words;
};

static PlanView plan_view(const std::vector<int> &plan, const TbxUnitHead &h)
{
    PlanView v;
    v.npair = h.counts & 255;
    v.nseg = (h.counts >> 8) & 255;
    v.nspair = h.counts >> 16;
    const int e1 = (10 * v.npair + 3) / 4, e2 = e1 + (v.npair + 3) / 4, e3 = e2 + v.nseg,
              e4 = e3 + (v.nspair + 3) / 4;
    const int *w = plan.data() + 4 * (size_t)h.plan;
    v.pairs = w;
    v.pext = w + 4 * e1;
    v.sinfo = w + 4 * e2;
    v.spl = w + 4 * e3;
    v.toff = w + 4 * e4;
    v.tsp = v.toff + 10;
    v.words = e4 + 5;
    return v;
}

extern "C" int tbx_batch_plan_stats(const tbx_batch_desc *desc, const tbx_tiles *tiles,
                                    const double *xyz, double skin, tbx_plan_stats *out)
{
    if (!desc || !out) return fail(TBX_ERR_ARG, "null argument");
    if (!(desc->cutoff > 0.0) || !std::isfinite(desc->cutoff))
        return fail(TBX_ERR_ARG, "invalid cutoff %g", desc->cutoff);
    HostSpecies hs;
    HostMols hm;
    int st = build_species(desc->bas, desc->h0, hs);
    if (!st) st = build_mols(hs, desc->nmol, desc->mol_nat, desc->id, hm);
    if (!st) st = set_hint(hm, xyz, skin, desc->cutoff);
    if (st) return st;
    const tbx_tiles t = tiles ? *tiles : kDefaultTiles;
    if ((st = check_tiles(hs, &t))) return st;
    if ((st = build_units(hs, hm, t))) return st;
    tbx_plan_stats s = {};
    s.ndense = hm.ndense;
    s.nunit_fwd = (long long)hm.fhead.size();
    s.nunit_grad = (long long)hm.uhead.size();
    s.plan_bytes_fwd = 4 * (long long)hm.fplan.size();
    s.plan_bytes_grad = 4 * (long long)hm.uplan.size();
    // atom pairs: one counter per ordered pair (i, j) of each molecule
    std::vector<long long> poff(hm.nmol + 1, 0);
    for (int m = 0; m < hm.nmol; ++m) {
        const long long n = hm.mol_atom[m + 1] - hm.mol_atom[m];
        poff[m + 1] = poff[m] + n * n;
    }
    for (int fwd = 0; fwd < 2; ++fwd) {
        const std::vector<TbxUnitHead> &head = fwd ? hm.fhead : hm.uhead;
        const std::vector<int> &plan = fwd ? hm.fplan : hm.uplan;
        const int cap_tasks = fwd ? t.fwd_unit_tasks : t.grad_unit_tasks;
        const int cap_pairs = fwd ? t.fwd_unit_pairs : t.grad_unit_pairs;
        std::vector<unsigned char> seen((size_t)poff[hm.nmol], 0);
        std::vector<unsigned char> hit(fwd ? (size_t)hm.ndense : 0, 0);
        for (const TbxUnitHead &h : head) {
            const PlanView v = plan_view(plan, h);
            const int ntask = v.toff[9];
            (fwd ? s.npair_fwd : s.npair_grad) += v.npair;
            (fwd ? s.ntask_fwd : s.ntask_grad) += ntask;
            if (ntask > cap_tasks || v.npair > cap_pairs || v.nseg > TBX_UNIT_SEGS ||
                v.nspair > ntask || (fwd && (h.nel & TBX_NEL_MASK) > 2 * t.fwd_unit_tasks))
                ++s.bad_units;
            bool live = false;
            for (int p = 0; p < v.npair; ++p) {
                const int i = v.pairs[10 * p], j = v.pairs[10 * p + 1];
                if (xyz && i >= 0 && i < hm.nat && j >= 0 && j < hm.nat && i != j) {
                    double r2 = 0.0;
                    for (int c = 0; c < 3; ++c) r2 += (xyz[3 * i + c] - xyz[3 * j + c]) *
                                                       (xyz[3 * i + c] - xyz[3 * j + c]);
                    live = live || r2 <= desc->cutoff * desc->cutoff;
                }
                const bool ok = i >= 0 && i < hm.nat && j >= 0 && j < hm.nat &&
                                hm.at_mol[i] == hm.at_mol[j] && (fwd ? j <= i : j < i) &&
                                h.mol == hm.at_mol[i];
                if (!ok) {
                    ++s.bad_pairs;
                    continue;
                }
                const int m = hm.at_mol[i], a0 = hm.mol_atom[m];
                const long long n = hm.mol_atom[m + 1] - a0;
                unsigned char &c = seen[poff[m] + (i - a0) * n + (j - a0)];
                if (c < 2) ++c;
            }
            if (!fwd && live) ++s.nunit_grad_live;
            if (!fwd) continue;
            const int m = h.mol;
            const long long nao = hm.mol_nao[m];
            const bool block_ok = h.dense0 == hm.mol_dense[m] && h.nao == nao;
            // the kernel's element phase: offsets and the locator from the
            // segment words, first task of each shell pair from the shell-pair list
            const int4 *seg = reinterpret_cast<const int4 *>(v.sinfo);
            const TbxPair *pinf = reinterpret_cast<const TbxPair *>(v.pairs);
            std::vector<int> off(2 * v.nseg + 1), roff(v.nseg + 1);
            tbx_elem_offsets(seg, v.nseg, off.data());
            // the row map the kernel reads the pair from (every row carries an
            // element, so it has at most 2 * TBX_UNIT_TASKS entries)
            tbx_row_offsets(seg, v.nseg, roff.data());
            std::vector<unsigned char> rmap(2 * TBX_UNIT_TASKS, 255);
            if (roff[v.nseg] > 2 * TBX_UNIT_TASKS) ++s.bad_units;
            else
                for (int q = 0; q < v.npair; ++q)
                    tbx_row_map_pair(seg, roff.data(), pinf, v.pext, v.npair, q, rmap.data());
            std::vector<int> spmap((size_t)16 * v.npair, -1);
            for (int q = 0; q < v.nspair; ++q) {
                const int e = v.spl[q], ty = e >> 16, li = ty / 3, lj = ty % 3;
                const int nc = (li + 1) * (li + 2) / 2 * ((lj + 1) * (lj + 2) / 2);
                const int pq = e & 255, jsh = (e >> 8) & 15, ish = (e >> 12) & 15;
                if (pq < v.npair && ish < 4 && jsh < 4)
                    spmap[16 * pq + 4 * ish + jsh] = v.toff[ty] + (q - v.tsp[ty]) * nc;
            }
            if (off[2 * v.nseg] != (h.nel & TBX_NEL_MASK)) ++s.bad_units;
            for (int e = 0; e < (h.nel & TBX_NEL_MASK); ++e) {
                ++s.nelem;
                const TbxElem el = tbx_elem_locate(off.data(), v.nseg, seg, pinf, v.pext, v.npair, e);
                const int p = el.p, nul = el.nul, muj = el.muj;
                const bool mirror = el.mirror != 0;
                bool map_ok = roff[v.nseg] <= 2 * TBX_UNIT_TASKS;
                if (map_ok) {
                    const TbxElem em = tbx_elem_locate_map(off.data(), v.nseg, seg, pinf,
                                                           roff.data(), rmap.data(), e);
                    map_ok = em.p == p && em.nul == nul && em.muj == muj && em.mirror == el.mirror;
                }
                if (!block_ok || !map_ok || p < 0 || p >= v.npair) {
                    ++s.bad_elems;
                    continue;
                }
                const int *pr = v.pairs + 10 * p;  // i, j, spi, spj, aoi, aoj, ...
                const int spi = pr[2], spj = pr[3];
                if (nul < 0 || muj < 0 || nul >= hs.nao[spi] || muj >= hs.nao[spj] ||
                    (mirror && pr[1] >= pr[0])) {
                    ++s.bad_elems;
                    continue;
                }
                // the element's Cartesian tasks are the block of its shell pair
                const int ish = hs.ao_sh[nul + TBX_MAXAO * spi];
                const int jsh = hs.ao_sh[muj + TBX_MAXAO * spj];
                const int li = hs.ang[ish + hs.msh * spi], lj = hs.ang[jsh + hs.msh * spj];
                const int ty = 3 * li + lj, nc = (li + 1) * (li + 2) / 2 * ((lj + 1) * (lj + 2) / 2);
                const int task0 = spmap[16 * p + 4 * ish + jsh];
                const int rel = task0 - v.toff[ty], spx = v.tsp[ty] + rel / nc;
                if (task0 < 0 || rel < 0 || rel % nc != 0 || task0 + nc > v.toff[ty + 1] ||
                    spx >= v.tsp[ty + 1] ||
                    v.spl[spx] != (p | (jsh << 8) | (ish << 12) | (ty << 16))) {
                    ++s.bad_elems;
                    continue;
                }
                const long long nu = pr[4] + nul, mu = pr[5] + muj;
                if (nu >= nao || mu >= nao) {
                    ++s.bad_elems;
                    continue;
                }
                unsigned char &c = hit[h.dense0 + (mirror ? mu * nao + nu : nu * nao + mu)];
                if (c < 2) ++c;
            }
        }
        for (int m = 0; m < hm.nmol; ++m) {
            const int a0 = hm.mol_atom[m];
            const long long n = hm.mol_atom[m + 1] - a0;
            for (long long i = 0; i < n; ++i)
                for (long long j = 0; j <= i; ++j) {
                    const int want = (fwd || j < i) ? 1 : 0;
                    if (seen[poff[m] + i * n + j] != want) ++s.bad_pairs;
                }
        }
        for (unsigned char c : hit)
            if (c != 1) ++s.bad_elems;
    }
    *out = s;
    return TBX_OK;
}

extern "C" void tbx_batch_delete(tbx_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    b->tab.release();
    b->flag.release();
    b->dse_tmp.release();
    b->rcov.release();
    b->sy_b.release();
    b->sy_e.release();
    b->sy_info.release();
    if (b->solver) (void)rocblas_destroy_handle(b->solver);
    for (auto &ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    delete b;
}

extern "C" int tbx_batch_sizes_get(const tbx_batch *b, tbx_batch_sizes *sz)
{
    if (!b || !sz) return fail(TBX_ERR_ARG, "null argument");
    sz->nmol = b->hm.nmol;
    sz->nat = b->hm.nat;
    sz->nsh = b->hm.nsh;
    sz->nao = b->hm.nao;
    sz->ndense = b->hm.ndense;
    sz->npair = 0;
    return TBX_OK;
}

extern "C" int tbx_batch_offsets(const tbx_batch *b, int *atom_off, int *sh_off, int *ao_off,
                                 long long *dense_off)
{
    if (!b) return fail(TBX_ERR_ARG, "null batch");
    const int n = b->hm.nmol + 1;
    if (atom_off) std::copy(b->hm.mol_atom.begin(), b->hm.mol_atom.end(), atom_off);
    if (sh_off) std::copy(b->hm.sh_off.begin(), b->hm.sh_off.end(), sh_off);
    if (ao_off) std::copy(b->hm.mol_ao.begin(), b->hm.mol_ao.end(), ao_off);
    if (dense_off) std::copy(b->hm.mol_dense.begin(), b->hm.mol_dense.end(), dense_off);
    (void)n;
    return TBX_OK;
}

// The caller's stream, as HIP defines it: NULL is the default (null) stream,
// which is also what torch reports for its default stream.
static hipStream_t pick_stream(tbx_batch *, void *stream)
{
    return static_cast<hipStream_t>(stream);
}

// Timing events are recorded only outside HIP graph capture.
static hipError_t mark(tbx_batch *b, int k, hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(s, &st);
    if (e != hipSuccess) return e;
    if (st != hipStreamCaptureStatusNone) return hipSuccess;
    return hipEventRecord(b->ev[k], s);
}

// The kernels move dipole and quadrupole components with 16-byte accesses.
static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

extern "C" int tbx_batch_forward(tbx_batch *b, const double *d_xyz, const double *d_se,
                                 double *d_S, double *d_D, double *d_Q, double *d_H, void *stream)
{
    if (!b || !d_xyz || !d_se || !d_S || !d_D || !d_Q || !d_H) return fail(TBX_ERR_ARG, "null argument");
    if (!aligned16(d_D) || !aligned16(d_Q))
        return fail(TBX_ERR_ARG, "d_D and d_Q must be 16-byte aligned");
    hipStream_t s = pick_stream(b, stream);
    HIP_TRY(mark(b, 0, s));
    HIP_TRY(tbx_launch_forward(&b->tab.sp, &b->tab.mt, d_xyz, d_se, d_S, d_D, d_Q, d_H,
                               b->hm.ndense, nullptr, b->flag.as<int>(), b->tiles.fwd_chunk, s));
    HIP_TRY(mark(b, 1, s));
    return TBX_OK;
}

static int batch_gradient(tbx_batch *b, const double *d_xyz, const double *d_se,
                          const double *d_S, const double *d_D, const double *d_dsedcn,
                          const double *d_P, const double *d_W, const double *d_vao,
                          const double *d_vdp, const double *d_vqp, double *d_dEdcn,
                          double *d_grad, double *d_sigma, void *stream)
{
    if (!b || !d_xyz || !d_se || !d_dsedcn || !d_P || !d_W || !d_vao || !d_vdp || !d_vqp ||
        !d_dEdcn || !d_grad || !d_sigma)
        return fail(TBX_ERR_ARG, "null argument");
    if (b->hs.diat) return fail(TBX_ERR_ARG, kNoDiatGradient);
    hipStream_t s = pick_stream(b, stream);
    HIP_TRY(mark(b, 4, s));
    HIP_TRY(tbx_launch_fill(d_grad, 3LL * b->hm.nat, 0.0, s));
    HIP_TRY(tbx_launch_fill(b->dse_tmp.as<double>(), b->hm.nsh, 0.0, s));
    HIP_TRY(tbx_launch_fill(d_sigma, 9LL * b->hm.nmol, 0.0, s));
    HIP_TRY(tbx_launch_gradient(0, &b->tab.sp, &b->tab.mt, d_xyz, d_se, d_P, d_W, d_vao, d_vdp,
                                d_vqp, 1, b->hm.ndense, b->hm.nao, nullptr, nullptr, nullptr,
                                nullptr, d_S, d_D, d_grad, b->dse_tmp.as<double>(), d_sigma,
                                nullptr, b->flag.as<int>(), b->tiles.grad_chunk, s));
    HIP_TRY(tbx_launch_dedcn(&b->tab.sp, &b->tab.mt, b->dse_tmp.as<double>(), d_dsedcn, d_P,
                             d_dEdcn, s));
    HIP_TRY(mark(b, 5, s));
    return TBX_OK;
}

extern "C" int tbx_batch_gradient(tbx_batch *b, const double *d_xyz, const double *d_se,
                                  const double *d_dsedcn, const double *d_P, const double *d_W,
                                  const double *d_vao, const double *d_vdp, const double *d_vqp,
                                  double *d_dEdcn, double *d_grad, double *d_sigma, void *stream)
{
    return batch_gradient(b, d_xyz, d_se, nullptr, nullptr, d_dsedcn, d_P, d_W, d_vao, d_vdp,
                          d_vqp, d_dEdcn, d_grad, d_sigma, stream);
}

extern "C" int tbx_batch_gradient_with_forward(tbx_batch *b, const double *d_xyz,
                                               const double *d_se, const double *d_S,
                                               const double *d_D, const double *d_dsedcn,
                                               const double *d_P, const double *d_W,
                                               const double *d_vao, const double *d_vdp,
                                               const double *d_vqp, double *d_dEdcn,
                                               double *d_grad, double *d_sigma, void *stream)
{
    if (!d_S || !d_D) return fail(TBX_ERR_ARG, "null argument");
    return batch_gradient(b, d_xyz, d_se, d_S, d_D, d_dsedcn, d_P, d_W, d_vao, d_vdp, d_vqp,
                          d_dEdcn, d_grad, d_sigma, stream);
}

// tbx_batch_gradient with the dE/dse partial sums in the static `scratch`
// (through its device address) instead of the batch's buffer: no device
// allocation and no allocating library call.
extern "C" int tbx_batch_gradient_noalloc(tbx_batch *b, const double *d_xyz, const double *d_se,
                                          const double *d_dsedcn, const double *d_P,
                                          const double *d_W, const double *d_vao,
                                          const double *d_vdp, const double *d_vqp,
                                          double *d_dEdcn, double *d_grad, double *d_sigma,
                                          void *stream)
{
    if (!b || !d_xyz || !d_se || !d_dsedcn || !d_P || !d_W || !d_vao || !d_vdp || !d_vqp ||
        !d_dEdcn || !d_grad || !d_sigma)
        return fail(TBX_ERR_ARG, "null argument");
    if (b->hs.diat) return fail(TBX_ERR_ARG, kNoDiatGradient);
    if (b->hm.nsh > TBX_NOALLOC_MAX_SHELLS)
        return fail(TBX_ERR_ARG,
                    "tbx_batch_gradient_noalloc: %d shells exceed the static scratch of %d; "
                    "use tbx_batch_gradient",
                    b->hm.nsh, TBX_NOALLOC_MAX_SHELLS);
    double *dse = scratch_ptr(b->ctx->device);
    if (!dse) return fail(TBX_ERR_DEVICE, "no static scratch on device %d", b->ctx->device);
    hipStream_t s = pick_stream(b, stream);
    // the scratch keeps the previous call's sums: clear it (a fill kernel)
    HIP_TRY(tbx_launch_fill(dse, b->hm.nsh, 0.0, s));
    HIP_TRY(tbx_launch_fill(d_grad, 3LL * b->hm.nat, 0.0, s));
    HIP_TRY(tbx_launch_fill(d_sigma, 9LL * b->hm.nmol, 0.0, s));
    HIP_TRY(tbx_launch_gradient(0, &b->tab.sp, &b->tab.mt, d_xyz, d_se, d_P, d_W, d_vao, d_vdp,
                                d_vqp, 1, b->hm.ndense, b->hm.nao, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, d_grad, dse, d_sigma, nullptr,
                                b->flag.as<int>(),
                                b->tiles.grad_chunk, s));
    HIP_TRY(tbx_launch_dedcn(&b->tab.sp, &b->tab.mt, dse, d_dsedcn, d_P, d_dEdcn, s));
    return TBX_OK;
}

static int batch_adjoint(tbx_batch *b, const double *d_xyz, const double *d_se,
                         const double *d_S, const double *d_D, const double *d_Sbar,
                         const double *d_Dbar, const double *d_Qbar, const double *d_Hbar,
                         double *d_dxyz, double *d_dse, void *stream)
{
    if (!b || !d_xyz || !d_se || !d_Sbar || !d_Dbar || !d_Qbar || !d_Hbar || !d_dxyz || !d_dse)
        return fail(TBX_ERR_ARG, "null argument");
    if (b->hs.diat) return fail(TBX_ERR_ARG, kNoDiatGradient);
    hipStream_t s = pick_stream(b, stream);
    HIP_TRY(mark(b, 6, s));
    HIP_TRY(tbx_launch_fill(d_dxyz, 3LL * b->hm.nat, 0.0, s));
    HIP_TRY(tbx_launch_fill(d_dse, b->hm.nsh, 0.0, s));
    HIP_TRY(tbx_launch_gradient(1, &b->tab.sp, &b->tab.mt, d_xyz, d_se, nullptr, nullptr,
                                nullptr, nullptr, nullptr, 1, 0, 0, d_Sbar, d_Dbar, d_Qbar,
                                d_Hbar, d_S, d_D, d_dxyz, d_dse, b->dse_tmp.as<double>(),
                                nullptr, b->flag.as<int>(), b->tiles.grad_chunk, s));
    HIP_TRY(mark(b, 7, s));
    return TBX_OK;
}

extern "C" int tbx_batch_adjoint(tbx_batch *b, const double *d_xyz, const double *d_se,
                                 const double *d_Sbar, const double *d_Dbar, const double *d_Qbar,
                                 const double *d_Hbar, double *d_dxyz, double *d_dse, void *stream)
{
    return batch_adjoint(b, d_xyz, d_se, nullptr, nullptr, d_Sbar, d_Dbar, d_Qbar, d_Hbar,
                         d_dxyz, d_dse, stream);
}

extern "C" int tbx_batch_adjoint_with_forward(tbx_batch *b, const double *d_xyz,
                                              const double *d_se, const double *d_S,
                                              const double *d_D, const double *d_Sbar,
                                              const double *d_Dbar, const double *d_Qbar,
                                              const double *d_Hbar, double *d_dxyz,
                                              double *d_dse, void *stream)
{
    if (!d_S || !d_D) return fail(TBX_ERR_ARG, "null argument");
    return batch_adjoint(b, d_xyz, d_se, d_S, d_D, d_Sbar, d_Dbar, d_Qbar, d_Hbar, d_dxyz,
                         d_dse, stream);
}

// Coordination number of the batch (mctc-lib ncoord "dexp" as GFN2 uses it,
// xtb/gfn2.f90:640, xtb/singlepoint.f90:224) and its chain rule.
extern "C" int tbx_batch_set_ncoord(tbx_batch *b, const double *rcov, double ka, double kb,
                                    double rshift, double cutoff)
{
    if (!b || !rcov) return fail(TBX_ERR_ARG, "null argument");
    for (double v : {ka, kb, cutoff})
        if (!(v > 0.0) || !std::isfinite(v))
            return fail(TBX_ERR_ARG, "ka, kb and cutoff must be positive and finite");
    if (!(rshift >= 0.0) || !std::isfinite(rshift)) return fail(TBX_ERR_ARG, "rshift %g", rshift);
    for (int k = 0; k < b->hs.nsp; ++k)
        if (!(rcov[k] > 0.0) || !std::isfinite(rcov[k]))
            return fail(TBX_ERR_ARG, "covalent radius of species %d is %g", k, rcov[k]);
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipDeviceSynchronize());  // no launch may still read the radii replaced here
    HIP_TRY(b->rcov.reserve(sizeof(double) * b->hs.nsp));
    HIP_TRY(hipMemcpy(b->rcov.p, rcov, sizeof(double) * b->hs.nsp, hipMemcpyHostToDevice));
    b->cn_ka = ka;
    b->cn_kb = kb;
    b->cn_rshift = rshift;
    b->cn_cutoff = cutoff;
    b->cn_set = true;
    return TBX_OK;
}

extern "C" int tbx_batch_ncoord(tbx_batch *b, const double *d_xyz, double *d_cn, void *stream)
{
    if (!b || !d_xyz || !d_cn) return fail(TBX_ERR_ARG, "null argument");
    if (!b->cn_set) return fail(TBX_ERR_ARG, "coordination number not set up (tbx_batch_set_ncoord)");
    HIP_TRY(tbx_launch_ncoord(&b->tab.mt, b->rcov.as<double>(), b->cn_ka, b->cn_kb, b->cn_rshift,
                              b->cn_cutoff, d_xyz, d_cn, pick_stream(b, stream)));
    return TBX_OK;
}

extern "C" int tbx_batch_ncoord_gradient(tbx_batch *b, const double *d_xyz,
                                         const double *d_dEdcn, double *d_grad, double *d_sigma,
                                         void *stream)
{
    if (!b || !d_xyz || !d_dEdcn || !d_grad || !d_sigma) return fail(TBX_ERR_ARG, "null argument");
    if (!b->cn_set) return fail(TBX_ERR_ARG, "coordination number not set up (tbx_batch_set_ncoord)");
    HIP_TRY(tbx_launch_ncoord_grad(&b->tab.mt, b->rcov.as<double>(), b->cn_ka, b->cn_kb,
                                   b->cn_rshift, b->cn_cutoff, d_xyz, d_dEdcn, d_grad, d_sigma,
                                   pick_stream(b, stream)));
    return TBX_OK;
}

// Density matrices of the SCF (scf/diag.f90:62-166, wavefunction/fermi.f90):
// the generalized eigenproblem per molecule (rocSOLVER dsygvd, as the
// reference's lapack/sygvd.f90 calls LAPACK's), the Fermi filling and the
// density products on the device.
extern "C" int tbx_batch_sygvd(tbx_batch *b, double *d_H, const double *d_S, double *d_eval,
                               void *stream)
{
    if (!b || !d_H || !d_S || !d_eval) return fail(TBX_ERR_ARG, "null argument");
    hipStream_t s = pick_stream(b, stream);
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (!b->solver && rocblas_create_handle(&b->solver) != rocblas_status_success)
        return fail(TBX_ERR_DEVICE, "rocblas_create_handle failed");
    if (rocblas_set_stream(b->solver, s) != rocblas_status_success)
        return fail(TBX_ERR_DEVICE, "rocblas_set_stream failed");
    // S is overwritten by its Cholesky factor: solve on a copy
    HIP_TRY(b->sy_b.reserve(sizeof(double) * (size_t)b->hm.ndense));
    HIP_TRY(b->sy_e.reserve(sizeof(double) * (size_t)std::max(1, b->hm.nao)));
    HIP_TRY(b->sy_info.reserve(sizeof(rocblas_int) * (size_t)b->hm.nmol));
    HIP_TRY(hipMemcpyAsync(b->sy_b.p, d_S, sizeof(double) * (size_t)b->hm.ndense,
                           hipMemcpyDeviceToDevice, s));
    for (int m = 0; m < b->hm.nmol; ++m) {
        const int n = b->hm.mol_nao[m];
        const long long d0 = b->hm.mol_dense[m];
        const int a0 = b->hm.mol_ao[m];
        const rocblas_status st = rocsolver_dsygvd(
            b->solver, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, d_H + d0,
            n, b->sy_b.as<double>() + d0, n, d_eval + a0, b->sy_e.as<double>() + a0,
            b->sy_info.as<rocblas_int>() + m);
        if (st != rocblas_status_success)
            return fail(TBX_ERR_DEVICE, "rocsolver_dsygvd (molecule %d): status %d", m, (int)st);
    }
    std::vector<rocblas_int> info(b->hm.nmol);
    HIP_TRY(hipMemcpyAsync(info.data(), b->sy_info.p, sizeof(rocblas_int) * info.size(),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int m = 0; m < b->hm.nmol; ++m)
        if (info[m] != 0)
            return fail(TBX_ERR_ARG, "sygvd failed for molecule %d (info %d): S not positive "
                                     "definite or no convergence", m, (int)info[m]);
    return TBX_OK;
}

extern "C" int tbx_batch_fermi_filling(tbx_batch *b, const double *d_nel, double kt,
                                       const double *d_emo, double *d_focc, double *d_efermi,
                                       void *stream)
{
    if (!b || !d_nel || !d_emo || !d_focc || !d_efermi) return fail(TBX_ERR_ARG, "null argument");
    if (!(kt > 0.0) || !std::isfinite(kt)) return fail(TBX_ERR_ARG, "kt %g", kt);
    HIP_TRY(tbx_launch_fermi(&b->tab.mt, d_nel, kt, d_emo, d_focc, d_efermi,
                             pick_stream(b, stream)));
    return TBX_OK;
}

extern "C" int tbx_batch_get_density_matrix(tbx_batch *b, const double *d_coeff,
                                            const double *d_focc, double *d_P, void *stream)
{
    if (!b || !d_coeff || !d_focc || !d_P) return fail(TBX_ERR_ARG, "null argument");
    HIP_TRY(tbx_launch_density(&b->tab.mt, d_coeff, d_focc, nullptr, d_P, pick_stream(b, stream)));
    return TBX_OK;
}

extern "C" int tbx_batch_get_wdensity(tbx_batch *b, const double *d_coeff, const double *d_focc,
                                      const double *d_eval, double *d_W, void *stream)
{
    if (!b || !d_coeff || !d_focc || !d_eval || !d_W) return fail(TBX_ERR_ARG, "null argument");
    HIP_TRY(tbx_launch_density(&b->tab.mt, d_coeff, d_focc, d_eval, d_W, pick_stream(b, stream)));
    return TBX_OK;
}

extern "C" int tbx_batch_last_timing(const tbx_batch *b, double *ms_forward, double *ms_pairs,
                                     double *ms_gradient, double *ms_adjoint)
{
    if (!b) return fail(TBX_ERR_ARG, "null batch");
    auto el = [&](int a, int c) {
        float ms = 0.0f;
        if (hipEventQuery(b->ev[c]) != hipSuccess) return -1.0;
        if (hipEventElapsedTime(&ms, b->ev[a], b->ev[c]) != hipSuccess) return -1.0;
        return (double)ms;
    };
    if (ms_forward) *ms_forward = el(0, 1);
    if (ms_pairs) *ms_pairs = 0.0;
    if (ms_gradient) *ms_gradient = el(4, 5);
    if (ms_adjoint) *ms_adjoint = el(6, 7);
    (void)hipGetLastError();  // a query of a never-recorded event is not an error here
    return TBX_OK;
}

// SCF-loop consumers of the resident integrals (scf/potential.f90:192-255,
// wavefunction/mulliken.f90:37-98): one launch each, events 8 + 2 what.
static int scf_launch(tbx_batch *b, int what, int nspin, int mpdim, int vdim, const double *S,
                      const double *mp, const double *P, const double *vao, const double *vmp,
                      const double *n0sh, double *h1, double *out, void *stream)
{
    if (nspin < 1 || nspin > 2) return fail(TBX_ERR_ARG, "nspin=%d", nspin);
    if (mpdim < 1 || mpdim > 6 || vdim < 1) return fail(TBX_ERR_ARG, "mpdim=%d vdim=%d", mpdim, vdim);
    if (mp && !aligned16(mp)) return fail(TBX_ERR_ARG, "d_mpint must be 16-byte aligned");
    hipStream_t s = pick_stream(b, stream);
    HIP_TRY(mark(b, 8 + 2 * what, s));
    HIP_TRY(tbx_launch_scf(what, &b->tab.mt, nspin, b->hm.ndense, mpdim, vdim, S, mp, P, vao, vmp,
                           n0sh, h1, out, s));
    HIP_TRY(mark(b, 9 + 2 * what, s));
    return TBX_OK;
}

extern "C" int tbx_batch_add_vao_to_h1(tbx_batch *b, int nspin, const double *d_S,
                                       const double *d_vao, double *d_h1, void *stream)
{
    if (!b || !d_S || !d_vao || !d_h1) return fail(TBX_ERR_ARG, "null argument");
    return scf_launch(b, 0, nspin, 1, 1, d_S, nullptr, nullptr, d_vao, nullptr, nullptr, d_h1,
                      nullptr, stream);
}

extern "C" int tbx_batch_add_vmp_to_h1(tbx_batch *b, int nspin, int mpdim, const double *d_mpint,
                                       int vdim, const double *d_vmp, double *d_h1, void *stream)
{
    if (!b || !d_mpint || !d_vmp || !d_h1) return fail(TBX_ERR_ARG, "null argument");
    return scf_launch(b, 1, nspin, mpdim, vdim, nullptr, d_mpint, nullptr, nullptr, d_vmp,
                      nullptr, d_h1, nullptr, stream);
}

extern "C" int tbx_batch_mulliken_shell_charges(tbx_batch *b, int nspin, const double *d_S,
                                                const double *d_P, const double *d_n0sh,
                                                double *d_qsh, void *stream)
{
    if (!b || !d_S || !d_P || !d_n0sh || !d_qsh) return fail(TBX_ERR_ARG, "null argument");
    return scf_launch(b, 2, nspin, 1, 1, d_S, nullptr, d_P, nullptr, nullptr, d_n0sh, nullptr,
                      d_qsh, stream);
}

extern "C" int tbx_batch_mulliken_atomic_multipoles(tbx_batch *b, int nspin, int mpdim,
                                                    const double *d_mpint, const double *d_P,
                                                    double *d_mpat, void *stream)
{
    if (!b || !d_mpint || !d_P || !d_mpat) return fail(TBX_ERR_ARG, "null argument");
    return scf_launch(b, 3, nspin, mpdim, 1, nullptr, d_mpint, d_P, nullptr, nullptr, nullptr,
                      nullptr, d_mpat, stream);
}

extern "C" int tbx_batch_last_scf_timing(const tbx_batch *b, double *ms)
{
    if (!b || !ms) return fail(TBX_ERR_ARG, "null argument");
    for (int k = 0; k < 4; ++k) {
        float t = 0.0f;
        ms[k] = -1.0;
        if (hipEventQuery(b->ev[9 + 2 * k]) == hipSuccess &&
            hipEventElapsedTime(&t, b->ev[8 + 2 * k], b->ev[9 + 2 * k]) == hipSuccess)
            ms[k] = t;
    }
    (void)hipGetLastError();  // a query of a never-recorded event is not an error here
    return TBX_OK;
}

extern "C" int tbx_probe_copy(const double *d_src, double *d_dst, long long n, void *stream)
{
    if (!d_src || !d_dst || n < 0) return fail(TBX_ERR_ARG, "tbx_probe_copy: bad arguments");
    HIP_TRY(tbx_launch_copy(d_src, d_dst, n, (hipStream_t)stream));
    return TBX_OK;
}
