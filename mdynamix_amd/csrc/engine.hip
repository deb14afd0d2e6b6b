// Host side of the MI355X force engine: the C ABI of include/mdx.h.
//
// One handle = one system on one HIP stream.  mdx_create copies the
// topology to HBM and precomputes everything static (k-vector rows as in
// INIFUR, forces.f:1356-1449; Ewald self energy; exclusion-correction
// pairs of ETERMS; the special-pair candidates).  mdx_build_list is CHCNB,
// mdx_compute is ALLFORCE / FFORCES / SLFORCES.  All device work of a call
// is queued on the handle's stream; the host waits only where a result
// must come back (list sizes after a rebuild, energies and forces).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "mdx_internal.h"

#include "list.hip"
#include "pair.hip"
#include "recip.hip"
#include "bonded.hip"

using namespace mdx;

namespace {

// Timed stages (HIP events on the handle's stream).
// Stage k >= 1 lies between events k-1 and k of compute_impl.
enum Stage { ST_LIST = 0, ST_PREP, ST_PAIRS, ST_SPECIAL, ST_RPREP, ST_SK, ST_SKRED, ST_RFORCE, ST_BONDED, ST_REDUCE, ST_N };

}  // namespace

struct mdx_engine {
    std::string err;
    hipStream_t stream = nullptr;
    // sizes
    int n = 0, nmol = 0, ntypes = 0, nlj = 0, ntypepairs = 0;
    mdx_params p{};
    Box box{};
    double fscale = 1.0;
    // device topology
    int *d_moltype = nullptr, *d_mol = nullptr, *d_ljtype = nullptr, *d_type_ee = nullptr;
    double *d_q = nullptr, *d_q0 = nullptr, *d_mass = nullptr, *d_c14el = nullptr, *d_c14lj = nullptr;
    double *d_a6_0 = nullptr, *d_b12_0 = nullptr, *d_a6_14 = nullptr, *d_b12_14 = nullptr;
    float2 *d_lj = nullptr;
    int *d_excl_start = nullptr, *d_excl_atom = nullptr;
    signed char *d_excl_is14 = nullptr;
    int *d_mstart = nullptr;
    std::vector<int> h_excl_start, h_excl_atom, h_mol, h_moltype, h_type_ee;
    std::vector<signed char> h_excl_is14;
    // specials and exclusion corrections
    int nspecial = 0;
    int2 *d_sp_pairs = nullptr;
    signed char *d_sp_kind = nullptr, *d_sp_cls = nullptr;
    std::vector<int2> h_sp_pairs;
    std::vector<signed char> h_sp_kind;
    int neterm = 0;
    int2 *d_et_pairs = nullptr;
    // bonded
    int nb = 0, na = 0, nt = 0;
    BondArgs bonds{};
    AngleArgs angles{};
    TorsArgs tors{};
    std::vector<void *> owned;
    // coordinates
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    bool have_pos = false;
    double *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr;
    bool user_com = false;
    // clusters and lists
    int ncx = 0, ncy = 0, ncol = 0, nclu = 0, nclu_cap = 0;
    unsigned long long *d_keys = nullptr, *d_keys2 = nullptr;
    int *d_vals = nullptr, *d_vals2 = nullptr;
    int *d_colcount = nullptr, *d_colstart = nullptr, *d_colclu_n = nullptr, *d_colclu = nullptr;
    void *d_cub = nullptr;
    size_t cub_bytes = 0;
    int *d_clu_atom = nullptr, *d_atom_slot = nullptr, *d_clu_col = nullptr;
    double *d_anchor = nullptr;
    float *d_ext = nullptr;
    int cand_cap = 0, list_cap = 0;
    int *d_cand = nullptr, *d_ncand = nullptr, *d_nlist = nullptr, *d_overflow = nullptr;
    PairEntry *d_list = nullptr;
    float4 *d_xq = nullptr;
    int2 *d_aux = nullptr;
    bool have_list = false;
    // reciprocal space
    int nkv = 0, nrows = 0, maxlen = 0, chunk = 512, nchunk = 0, nslice = 1, rows_per_slice = 0;
    double kbox[3] = {0, 0, 0};
    double spei = 0.0;
    std::vector<double> spee_self;
    RecipRow *d_rows = nullptr;
    int *d_rowoff = nullptr;
    int4 *d_kidx = nullptr;
    double4 *d_fq = nullptr;
    double4 *d_stepz = nullptr;
    float2 *d_part = nullptr, *d_ak = nullptr;
    int *d_wavelen = nullptr;
    double2 *d_sk = nullptr;
    double *d_rpart = nullptr;
    std::vector<int4> h_kidx;   // (nx, ny, nz) in INIFUR order
    // forces
    double *d_fs = nullptr, *d_fl = nullptr;     // [3][n] each
    unsigned long long *d_fix = nullptr;          // [9][n]: G, HL, HB
    double *d_out = nullptr;                      // [9][n]: F, G, H
    double *d_epart = nullptr;
    int epart_rows = 0;
    double *d_energy = nullptr;
    // timing
    hipEvent_t ev[12] = {};
    double stage_ms[ST_N] = {};
    bool timing = true;
};

#define HIPCHK(h, call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            (h)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                       #call;                                                        \
            return MDX_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

#define CHK(call)                    \
    do {                             \
        mdx_status s_ = (call);      \
        if (s_ != MDX_OK) return s_; \
    } while (0)

namespace {

template <typename T>
mdx_status dalloc(mdx_engine *h, T **p, size_t count) {
    if (*p) { hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        h->err = "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed";
        return MDX_ERR_NOMEM;
    }
    return MDX_OK;
}

template <typename T>
mdx_status upload(mdx_engine *h, T **p, const T *src, size_t count) {
    *p = nullptr;
    CHK(dalloc(h, p, count));
    if (count && src) HIPCHK(h, hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyHostToDevice));
    h->owned.push_back((void *)*p);
    return MDX_OK;
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// INIFUR enumeration (forces.f:1365-1427), rectangular cell.
void enumerate_k(const mdx_params &p, std::vector<int4> &out) {
    out.clear();
    const double alp2 = p.alpha * p.alpha;
    const double cutk2 = 4.0 * alp2 * p.fexp;
    const double cutk = std::sqrt(cutk2);
    const double twopi = 2.0 * kPI;
    const int kmax = (int)(cutk * p.box[0] / twopi + 1.0);
    const int kmay = (int)(cutk * p.box[1] / twopi + 1.0);
    const int kmaz = (int)(cutk * p.box[2] / twopi + 1.0);
    for (int ikz = -kmaz; ikz <= kmaz; ++ikz)
        for (int ikx = 0; ikx <= kmax; ++ikx) {
            const int ky1 = ikx == 0 ? (ikz > 0 ? 0 : 1) : -kmay;
            for (int iky = ky1; iky <= kmay; ++iky) {
                const double ax = twopi * ikx / p.box[0], ay = twopi * iky / p.box[1], az = twopi * ikz / p.box[2];
                double rk2 = ax * ax + ay * ay;
                rk2 = rk2 + az * az;
                if (rk2 <= cutk2) out.push_back(make_int4(ikx, iky, ikz, 0));
            }
        }
}

mdx_status setup_kspace(mdx_engine *h) {
    std::vector<int4> ks;
    enumerate_k(h->p, ks);
    h->nkv = (int)ks.size();
    if (h->nkv > 4000000) {
        h->err = "too many k-vectors";
        return MDX_ERR_NKV;
    }
    h->h_kidx = ks;
    for (int k = 0; k < 3; ++k) h->kbox[k] = h->p.box[k];
    // rows of fixed (nx, ny)
    std::map<std::pair<int, int>, std::pair<int, int>> rng;
    for (const int4 &k : ks) {
        auto key = std::make_pair(k.x, k.y);
        auto it = rng.find(key);
        if (it == rng.end()) rng[key] = std::make_pair(k.z, k.z);
        else {
            it->second.first = std::min(it->second.first, k.z);
            it->second.second = std::max(it->second.second, k.z);
        }
    }
    std::vector<RecipRow> rows;
    for (auto &kv : rng) {
        RecipRow r;
        r.nx = kv.first.first;
        r.ny = kv.first.second;
        r.nz0 = kv.second.first;
        r.len = kv.second.second - kv.second.first + 1;
        rows.push_back(r);
    }
    std::stable_sort(rows.begin(), rows.end(), [](const RecipRow &a, const RecipRow &b) { return a.len > b.len; });
    long long total = 0;
    std::vector<int> off(rows.size());
    std::vector<int4> kidx;
    int maxlen = 0;
    for (size_t r = 0; r < rows.size(); ++r) {
        off[r] = (int)total;
        for (int j = 0; j < rows[r].len; ++j) kidx.push_back(make_int4(rows[r].nx, rows[r].ny, rows[r].nz0 + j, 0));
        total += rows[r].len;
        maxlen = std::max(maxlen, rows[r].len);
    }
    if (total != h->nkv) {
        h->err = "k-vector rows are not contiguous";
        return MDX_ERR_NKV;
    }
    if (maxlen > 96) {
        h->err = "k-vector rows longer than 96 (kmax > 47) are not supported";
        return MDX_ERR_NKV;
    }
    h->nrows = (int)rows.size();
    h->maxlen = maxlen;
    // longest row of each wavefront of 64 rows (rows are sorted by length)
    std::vector<int> wl((rows.size() + 63) / 64, 0);
    for (size_t r = 0; r < rows.size(); ++r) wl[r / 64] = std::max(wl[r / 64], rows[r].len);
    CHK(dalloc(h, &h->d_wavelen, wl.size()));
    HIPCHK(h, hipMemcpy(h->d_wavelen, wl.data(), sizeof(int) * wl.size(), hipMemcpyHostToDevice));
    CHK(dalloc(h, &h->d_rows, rows.size()));
    CHK(dalloc(h, &h->d_rowoff, rows.size()));
    CHK(dalloc(h, &h->d_kidx, kidx.size()));
    HIPCHK(h, hipMemcpy(h->d_rows, rows.data(), sizeof(RecipRow) * rows.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_rowoff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_kidx, kidx.data(), sizeof(int4) * kidx.size(), hipMemcpyHostToDevice));
    CHK(dalloc(h, &h->d_fq, h->n));
    CHK(dalloc(h, &h->d_stepz, h->n));
    h->nchunk = cdiv(h->n, h->chunk);
    CHK(dalloc(h, &h->d_part, (size_t)h->nchunk * std::max(1, h->nkv)));
    CHK(dalloc(h, &h->d_ak, std::max(1, h->nkv)));
    CHK(dalloc(h, &h->d_sk, std::max(1, h->nkv)));
    // force pass: slices of rows so the grid fills the chip
    const int wav = cdiv(h->n, 64);
    h->nslice = std::max(1, std::min(16, cdiv(4096, std::max(1, wav))));
    h->rows_per_slice = cdiv(std::max(1, h->nrows), h->nslice);
    CHK(dalloc(h, &h->d_rpart, (size_t)h->nslice * 3 * h->n));
    return MDX_OK;
}

void self_energy(mdx_engine *h, const mdx_topology *t) {
    // forces.f:1429-1447: SPEEI(ITYP) = -NSPEC * alpha/sqrt(pi) * COULF * sum (ESCI q)^2
    const double abc = h->p.alpha / std::sqrt(kPI);
    h->spee_self.assign(h->ntypes, 0.0);
    h->spei = 0.0;
    for (int it = 0; it < h->ntypes; ++it) {
        const double esci = t->type_ee[it] == 0 ? h->p.ee_self_scale : 1.0;
        int first = -1, nsits = 0, nspec = 0, last = -1;
        for (int i = 0; i < h->n; ++i)
            if (t->atom_moltype[i] == it) {
                if (first < 0) first = i;
                if (t->atom_mol[i] == t->atom_mol[first]) ++nsits;
                if (t->atom_mol[i] != last) { ++nspec; last = t->atom_mol[i]; }
            }
        double spet = 0.0;
        for (int s = 0; first >= 0 && s < nsits; ++s) {
            const double qs = esci * t->atom_charge0[first + s];
            spet = spet - (double)nspec * abc * (h->p.coulomb_factor * (qs * qs));
        }
        h->spee_self[it] = spet;
        h->spei = h->spei + spet;
    }
}

mdx_status alloc_clusters(mdx_engine *h) {
    const double vol = h->p.box[0] * h->p.box[1] * h->p.box[2];
    const double rho = h->n / vol;
    const double cw = std::cbrt(kClu / rho);
    h->ncx = std::max(1, (int)(h->p.box[0] / cw));
    h->ncy = std::max(1, (int)(h->p.box[1] / cw));
    h->ncol = h->ncx * h->ncy;
    h->nclu_cap = h->n / kClu + h->ncol + 1;
    CHK(dalloc(h, &h->d_keys, h->n));
    CHK(dalloc(h, &h->d_keys2, h->n));
    CHK(dalloc(h, &h->d_vals, h->n));
    CHK(dalloc(h, &h->d_vals2, h->n));
    CHK(dalloc(h, &h->d_colcount, h->ncol));
    CHK(dalloc(h, &h->d_colstart, h->ncol));
    CHK(dalloc(h, &h->d_colclu_n, h->ncol));
    CHK(dalloc(h, &h->d_colclu, h->ncol));
    CHK(dalloc(h, &h->d_clu_atom, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_atom_slot, h->n));
    CHK(dalloc(h, &h->d_clu_col, h->nclu_cap));
    CHK(dalloc(h, &h->d_anchor, (size_t)h->nclu_cap * 3));
    CHK(dalloc(h, &h->d_ext, (size_t)h->nclu_cap * 3));
    CHK(dalloc(h, &h->d_ncand, h->nclu_cap));
    CHK(dalloc(h, &h->d_nlist, h->nclu_cap));
    CHK(dalloc(h, &h->d_overflow, 1));
    CHK(dalloc(h, &h->d_xq, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_aux, (size_t)h->nclu_cap * kClu));
    // temporary storage of the sort and scans
    size_t b1 = 0, b2 = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, b1, h->d_keys, h->d_keys2, h->d_vals, h->d_vals2, h->n, 0, 64);
    hipcub::DeviceScan::ExclusiveSum(nullptr, b2, h->d_colcount, h->d_colstart, h->ncol);
    h->cub_bytes = std::max(b1, b2);
    CHK(dalloc(h, (char **)&h->d_cub, h->cub_bytes));
    // list capacities from the density: clusters whose boxes come within
    // r_list + 2 cluster extents
    const double rl = h->p.rcut + (h->p.list_mode == MDX_LIST_VERLET ? h->p.skin : 0.0);
    const double reach = rl + 2.0 * cw;
    double est = 4.0 / 3.0 * kPI * reach * reach * reach * rho / kClu * 1.3 + 32;
    est = std::min(est, (double)h->nclu_cap);
    h->cand_cap = ((int)est + 63) / 64 * 64;
    h->list_cap = h->cand_cap;
    CHK(dalloc(h, &h->d_cand, (size_t)h->nclu_cap * h->cand_cap));
    CHK(dalloc(h, &h->d_list, (size_t)h->nclu_cap * h->list_cap));
    // per-call buffers
    CHK(dalloc(h, &h->d_fs, (size_t)3 * h->n));
    CHK(dalloc(h, &h->d_fl, (size_t)3 * h->n));
    CHK(dalloc(h, &h->d_fix, (size_t)9 * h->n));
    CHK(dalloc(h, &h->d_out, (size_t)9 * h->n));
    return MDX_OK;
}

mdx_status ensure_epart(mdx_engine *h, int rows) {
    if (rows > h->epart_rows) {
        CHK(dalloc(h, &h->d_epart, (size_t)rows * Q_N));
        h->epart_rows = rows;
    }
    return MDX_OK;
}

double list_radius2(const mdx_engine *h) {
    const double rl = h->p.rcut + (h->p.list_mode == MDX_LIST_VERLET ? h->p.skin : 0.0);
    return rl * rl;
}

}  // namespace

extern "C" {

const char *mdx_version(void) { return "mdx-mi355x 0.1 (gfx950)"; }

mdx_status mdx_device_info(int32_t *n_cu, int32_t *clock_mhz, char *name, int32_t name_len) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return MDX_ERR_HIP;
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_mhz) *clock_mhz = prop.clockRate / 1000;
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return MDX_OK;
}

const char *mdx_last_error(const mdx_engine *h) { return h ? h->err.c_str() : "null handle"; }

mdx_status mdx_create(const mdx_topology *t, const mdx_params *p, mdx_engine **out) {
    if (!t || !p || !out) return MDX_ERR_ARG;
    *out = nullptr;
    mdx_engine *h = new mdx_engine();
    auto fail = [&](mdx_status s) { *out = h; return s; };
    h->n = t->natoms;
    h->nmol = t->nmols;
    h->ntypes = t->nmoltypes;
    h->nlj = t->nljtypes;
    h->ntypepairs = t->ntypepairs;
    h->p = *p;
    if (h->n <= 0 || h->ntypes <= 0 || h->ntypes > MDX_MAX_MOLTYPES || h->nlj <= 0 || h->nlj > 32) {
        h->err = "unsupported sizes (1..16 molecule types, 1..32 LJ types)";
        return fail(MDX_ERR_ARG);
    }
    for (int i = 0; i < h->ntypes; ++i)
        if (!t->type_move[i]) {
            h->err = "fixed molecules (LMOVE = .false.) are not supported";
            return fail(MDX_ERR_ARG);
        }
    {
        const int n = h->n;
        auto bad = [&](const char *what) { h->err = std::string("index out of range: ") + what; return fail(MDX_ERR_ARG); };
        for (int i = 0; i < n; ++i) {
            if (t->atom_moltype[i] < 0 || t->atom_moltype[i] >= h->ntypes) return bad("atom_moltype");
            if (t->atom_ljtype[i] < 0 || t->atom_ljtype[i] >= h->nlj) return bad("atom_ljtype");
            if (t->excl_start[i + 1] < t->excl_start[i]) return bad("excl_start");
        }
        for (int k = 0; k < t->excl_start[n]; ++k)
            if (t->excl_atom[k] < 0 || t->excl_atom[k] >= n) return bad("excl_atom");
        for (int k = 0; k < h->ntypes * h->ntypes; ++k)
            if (t->type_pair[k] < 0 || t->type_pair[k] >= MDX_MAX_TYPEPAIRS) return bad("type_pair");
        for (int k = 0; k < t->nbonds; ++k)
            if (t->bond_i[k] < 0 || t->bond_i[k] >= n || t->bond_j[k] < 0 || t->bond_j[k] >= n ||
                t->bond_param[k] < 0 || t->bond_param[k] >= t->nbondparams) return bad("bond");
        for (int k = 0; k < t->nangles; ++k)
            if (t->angle_i[k] < 0 || t->angle_i[k] >= n || t->angle_j[k] < 0 || t->angle_j[k] >= n ||
                t->angle_k[k] < 0 || t->angle_k[k] >= n || t->angle_param[k] < 0 ||
                t->angle_param[k] >= t->nangleparams) return bad("angle");
        for (int k = 0; k < t->ntors; ++k)
            if (t->tors_i[k] < 0 || t->tors_i[k] >= n || t->tors_j[k] < 0 || t->tors_j[k] >= n ||
                t->tors_k[k] < 0 || t->tors_k[k] >= n || t->tors_l[k] < 0 || t->tors_l[k] >= n ||
                t->tors_param[k] < 0 || t->tors_param[k] >= t->ntorsparams) return bad("dihedral");
        if (!(p->coulomb_factor > 0.0)) { h->err = "coulomb_factor must be positive"; return fail(MDX_ERR_ARG); }
    }
    if (!(p->alpha > 0.0)) {
        h->err = "only Ewald electrostatics (alpha > 0) is supported";
        return fail(MDX_ERR_ARG);
    }
    const double rl = p->rcut + (p->list_mode == MDX_LIST_VERLET ? p->skin : 0.0);
    for (int k = 0; k < 3; ++k) {
        h->box.l[k] = p->box[k];
        h->box.h[k] = 0.5 * p->box[k];
        if (!(rl < 0.5 * p->box[k])) {
            h->err = "list radius must be below half the box edge";
            return fail(MDX_ERR_ARG);
        }
    }
    if (hipStreamCreate(&h->stream) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return fail(MDX_ERR_HIP);
    }
    for (int k = 0; k < 12; ++k) hipEventCreate(&h->ev[k]);
    const int n = h->n;
    // fixed-point scale: resolution ~1e-14 of coulomb_factor, range ~1e4 * coulomb_factor
    h->fscale = std::ldexp(1.0, (int)std::floor(std::log2(std::ldexp(1.0, 60) / (std::fabs(p->coulomb_factor) * 1e4 + 1e-300))));
    // topology
    h->h_mol.assign(t->atom_mol, t->atom_mol + n);
    h->h_moltype.assign(t->atom_moltype, t->atom_moltype + n);
    h->h_type_ee.assign(t->type_ee, t->type_ee + h->ntypes);
    h->h_excl_start.assign(t->excl_start, t->excl_start + n + 1);
    const int nex = h->h_excl_start[n];
    h->h_excl_atom.assign(t->excl_atom, t->excl_atom + nex);
    h->h_excl_is14.assign(t->excl_is14, t->excl_is14 + nex);
    mdx_status s;
    if ((s = upload(h, &h->d_moltype, t->atom_moltype, n)) ||
        (s = upload(h, &h->d_mol, t->atom_mol, n)) || (s = upload(h, &h->d_ljtype, t->atom_ljtype, n)) ||
        (s = upload(h, &h->d_type_ee, t->type_ee, h->ntypes)) || (s = upload(h, &h->d_q, t->atom_charge, n)) ||
        (s = upload(h, &h->d_q0, t->atom_charge0, n)) ||
        (s = upload(h, &h->d_c14el, t->type_c14el, h->ntypes)) ||
        (s = upload(h, &h->d_c14lj, t->type_c14lj, h->ntypes)) ||
        (s = upload(h, &h->d_a6_0, t->a6_0, (size_t)h->nlj * h->nlj)) ||
        (s = upload(h, &h->d_b12_0, t->b12_0, (size_t)h->nlj * h->nlj)) ||
        (s = upload(h, &h->d_a6_14, t->a6_14, (size_t)h->nlj * h->nlj)) ||
        (s = upload(h, &h->d_b12_14, t->b12_14, (size_t)h->nlj * h->nlj)) ||
        (s = upload(h, &h->d_excl_start, t->excl_start, n + 1)) ||
        (s = upload(h, &h->d_excl_atom, t->excl_atom, nex)) ||
        (s = upload(h, &h->d_excl_is14, (const signed char *)t->excl_is14, nex)))
        return fail(s);
    std::vector<double> mass(n, 1.0);
    if (t->atom_mass) mass.assign(t->atom_mass, t->atom_mass + n);
    if ((s = upload(h, &h->d_mass, mass.data(), n))) return fail(s);
    std::vector<float2> lj((size_t)h->nlj * h->nlj);
    for (size_t k = 0; k < lj.size(); ++k) lj[k] = make_float2((float)t->a6[k], (float)t->b12[k]);
    if ((s = upload(h, &h->d_lj, lj.data(), lj.size()))) return fail(s);
    // molecules are contiguous atom ranges
    std::vector<int> mstart(h->nmol + 1, 0);
    for (int i = 0; i < n; ++i) {
        const int m = t->atom_mol[i];
        if (m < 0 || m >= h->nmol || (i > 0 && m != t->atom_mol[i - 1] && m != t->atom_mol[i - 1] + 1)) {
            h->err = "atoms must be ordered molecule by molecule";
            return fail(MDX_ERR_ARG);
        }
        mstart[m + 1] = i + 1;
    }
    if ((s = upload(h, &h->d_mstart, mstart.data(), h->nmol + 1))) return fail(s);
    // special-pair candidates and exclusion-correction pairs
    std::vector<int2> et;
    for (int i = 0; i < n; ++i) {
        const int it = t->atom_moltype[i];
        if (t->type_ee[it] != 0) {
            for (int k = h->h_excl_start[i]; k < h->h_excl_start[i + 1]; ++k) {
                const int j = h->h_excl_atom[k];
                if (i > j) {
                    et.push_back(make_int2(i, j));
                    if (h->h_excl_is14[k]) { h->h_sp_pairs.push_back(make_int2(i, j)); h->h_sp_kind.push_back(2); }
                }
            }
        }
    }
    for (int m = 0; m < h->nmol; ++m) {
        const int a0 = mstart[m], a1 = mstart[m + 1];
        if (a1 <= a0 || t->type_ee[t->atom_moltype[a0]] != 0) continue;
        for (int i = a0 + 1; i < a1; ++i)
            for (int j = a0; j < i; ++j) {
                et.push_back(make_int2(i, j));
                int kind = 4;
                for (int k = h->h_excl_start[i]; k < h->h_excl_start[i + 1]; ++k)
                    if (h->h_excl_atom[k] == j) kind |= h->h_excl_is14[k] ? 2 : 1;
                h->h_sp_pairs.push_back(make_int2(i, j));
                h->h_sp_kind.push_back((signed char)kind);
            }
    }
    h->neterm = (int)et.size();
    h->nspecial = (int)h->h_sp_pairs.size();
    if ((s = upload(h, &h->d_et_pairs, et.data(), et.size())) ||
        (s = upload(h, &h->d_sp_pairs, h->h_sp_pairs.data(), h->h_sp_pairs.size())) ||
        (s = upload(h, &h->d_sp_kind, h->h_sp_kind.data(), h->h_sp_kind.size())))
        return fail(s);
    if ((s = dalloc(h, &h->d_sp_cls, std::max(1, h->nspecial)))) return fail(s);
    // bonded terms
    h->nb = t->nbonds;
    h->na = t->nangles;
    h->nt = t->ntors;
    int *ip;
    double *dp;
    h->bonds.n = h->nb;
    if ((s = upload(h, &ip, t->bond_i, h->nb))) return fail(s); h->bonds.bi = ip;
    if ((s = upload(h, &ip, t->bond_j, h->nb))) return fail(s); h->bonds.bj = ip;
    if ((s = upload(h, &ip, t->bond_param, h->nb))) return fail(s); h->bonds.bpar = ip;
    if ((s = upload(h, &ip, t->bp_morse, t->nbondparams))) return fail(s); h->bonds.morse = ip;
    if ((s = upload(h, &dp, t->bp_r0, t->nbondparams))) return fail(s); h->bonds.r0 = dp;
    if ((s = upload(h, &dp, t->bp_k, t->nbondparams))) return fail(s); h->bonds.k = dp;
    if ((s = upload(h, &dp, t->bp_d, t->nbondparams))) return fail(s); h->bonds.d = dp;
    if ((s = upload(h, &dp, t->bp_rho, t->nbondparams))) return fail(s); h->bonds.rho = dp;
    h->angles.n = h->na;
    if ((s = upload(h, &ip, t->angle_i, h->na))) return fail(s); h->angles.ai = ip;
    if ((s = upload(h, &ip, t->angle_j, h->na))) return fail(s); h->angles.aj = ip;
    if ((s = upload(h, &ip, t->angle_k, h->na))) return fail(s); h->angles.ak = ip;
    if ((s = upload(h, &ip, t->angle_param, h->na))) return fail(s); h->angles.apar = ip;
    if ((s = upload(h, &dp, t->ap_theta0, t->nangleparams))) return fail(s); h->angles.theta0 = dp;
    if ((s = upload(h, &dp, t->ap_k, t->nangleparams))) return fail(s); h->angles.k = dp;
    h->tors.n = h->nt;
    if ((s = upload(h, &ip, t->tors_i, h->nt))) return fail(s); h->tors.ti = ip;
    if ((s = upload(h, &ip, t->tors_j, h->nt))) return fail(s); h->tors.tj = ip;
    if ((s = upload(h, &ip, t->tors_k, h->nt))) return fail(s); h->tors.tk = ip;
    if ((s = upload(h, &ip, t->tors_l, h->nt))) return fail(s); h->tors.tl = ip;
    if ((s = upload(h, &ip, t->tors_param, h->nt))) return fail(s); h->tors.tpar = ip;
    if ((s = upload(h, &ip, t->tp_kind, t->ntorsparams))) return fail(s); h->tors.kind = ip;
    if ((s = upload(h, &ip, t->tp_mult, t->ntorsparams))) return fail(s); h->tors.mult = ip;
    if ((s = upload(h, &dp, t->tp_k, t->ntorsparams))) return fail(s); h->tors.k = dp;
    if ((s = upload(h, &dp, t->tp_delta, t->ntorsparams))) return fail(s); h->tors.delta = dp;
    if ((s = upload(h, &dp, t->tp_k1, t->ntorsparams))) return fail(s); h->tors.k1 = dp;
    if ((s = upload(h, &dp, t->tp_k2, t->ntorsparams))) return fail(s); h->tors.k2 = dp;
    if ((s = upload(h, &dp, t->tp_k3, t->ntorsparams))) return fail(s); h->tors.k3 = dp;
    if ((s = upload(h, &dp, t->tp_k4, t->ntorsparams))) return fail(s); h->tors.k4 = dp;
    if ((s = upload(h, &dp, t->tp_k5, t->ntorsparams))) return fail(s); h->tors.k5 = dp;
    // coordinates, COMs, energies
    if ((s = dalloc(h, &h->d_x, n)) || (s = dalloc(h, &h->d_y, n)) || (s = dalloc(h, &h->d_z, n)) ||
        (s = dalloc(h, &h->d_cx, h->nmol)) || (s = dalloc(h, &h->d_cy, h->nmol)) ||
        (s = dalloc(h, &h->d_cz, h->nmol)) || (s = dalloc(h, &h->d_energy, Q_N)))
        return fail(s);
    self_energy(h, t);
    if ((s = setup_kspace(h))) return fail(s);
    if ((s = alloc_clusters(h))) return fail(s);
    *out = h;
    return MDX_OK;
}

mdx_status mdx_destroy(mdx_engine *h) {
    if (!h) return MDX_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void *q : h->owned) hipFree(q);
    void *bufs[] = {h->d_sp_cls, h->d_x, h->d_y, h->d_z, h->d_cx, h->d_cy, h->d_cz, h->d_keys, h->d_keys2,
                    h->d_vals, h->d_vals2, h->d_colcount, h->d_colstart, h->d_colclu_n, h->d_colclu, h->d_cub,
                    h->d_clu_atom, h->d_atom_slot, h->d_clu_col, h->d_anchor, h->d_ext, h->d_cand, h->d_ncand,
                    h->d_nlist, h->d_overflow, h->d_list, h->d_xq, h->d_aux, h->d_rows, h->d_rowoff, h->d_kidx,
                    h->d_fq, h->d_stepz, h->d_part, h->d_ak, h->d_sk, h->d_rpart, h->d_wavelen, h->d_fs, h->d_fl, h->d_fix,
                    h->d_out, h->d_epart, h->d_energy};
    for (void *q : bufs)
        if (q) hipFree(q);
    for (int k = 0; k < 12; ++k)
        if (h->ev[k]) hipEventDestroy(h->ev[k]);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return MDX_OK;
}

mdx_status mdx_set_positions(mdx_engine *h, const double *x, const double *y, const double *z) {
    if (!h || !x || !y || !z) return MDX_ERR_ARG;
    const size_t b = sizeof(double) * h->n;
    HIPCHK(h, hipMemcpyAsync(h->d_x, x, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_y, y, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_z, z, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_pos = true;
    return MDX_OK;
}

mdx_status mdx_set_positions_device(mdx_engine *h, const double *x, const double *y, const double *z) {
    if (!h || !x || !y || !z) return MDX_ERR_ARG;
    const size_t b = sizeof(double) * h->n;
    HIPCHK(h, hipMemcpyAsync(h->d_x, x, b, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_y, y, b, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_z, z, b, hipMemcpyDeviceToDevice, h->stream));
    h->have_pos = true;
    return MDX_OK;
}

mdx_status mdx_set_molecule_com(mdx_engine *h, const double *x, const double *y, const double *z) {
    if (!h) return MDX_ERR_ARG;
    if (!x || !y || !z) {
        h->user_com = false;
        return MDX_OK;
    }
    const size_t b = sizeof(double) * h->nmol;
    HIPCHK(h, hipMemcpyAsync(h->d_cx, x, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_cy, y, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_cz, z, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->user_com = true;
    return MDX_OK;
}

mdx_status mdx_set_box(mdx_engine *h, const double box[3]) {
    if (!h || !box) return MDX_ERR_ARG;
    const double rl = h->p.rcut + (h->p.list_mode == MDX_LIST_VERLET ? h->p.skin : 0.0);
    for (int k = 0; k < 3; ++k)
        if (!(rl < 0.5 * box[k])) {
            h->err = "list radius must be below half the box edge";
            return MDX_ERR_ARG;
        }
    bool redo = false;
    for (int k = 0; k < 3; ++k) {
        h->p.box[k] = box[k];
        h->box.l[k] = box[k];
        h->box.h[k] = 0.5 * box[k];
        // FURIR re-runs INIFUR when an edge changed by more than 5 % (forces.f:1251-1259;
        // FKTX/FKTS are REAL*4 literals)
        const double r = box[k] / h->kbox[k];
        if (r > (double)1.05f || r < (double)0.95f) redo = true;
    }
    if (redo) CHK(setup_kspace(h));
    h->have_list = false;
    return MDX_OK;
}

mdx_status mdx_build_list(mdx_engine *h) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_pos) {
        h->err = "mdx_build_list before mdx_set_positions";
        return MDX_ERR_STATE;
    }
    hipStream_t st = h->stream;
    const int n = h->n;
    if (h->timing) hipEventRecord(h->ev[0], st);
    k_fill_i32<<<cdiv(h->ncol, 256), 256, 0, st>>>(h->ncol, h->d_colcount, 0);
    k_bin<<<cdiv(n, 256), 256, 0, st>>>(n, h->d_x, h->d_y, h->d_z, h->box, h->ncx, h->ncy, h->d_keys, h->d_vals,
                                         h->d_colcount);
    int colbits = 1;
    while ((1 << colbits) < h->ncol) ++colbits;
    size_t tb = h->cub_bytes;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(h->d_cub, tb, h->d_keys, h->d_keys2, h->d_vals, h->d_vals2, n, 0,
                                                 32 + colbits, st));
    tb = h->cub_bytes;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->d_cub, tb, h->d_colcount, h->d_colstart, h->ncol, st));
    k_col_clusters<<<cdiv(h->ncol, 256), 256, 0, st>>>(h->ncol, h->d_colcount, h->d_colclu_n);
    tb = h->cub_bytes;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->d_cub, tb, h->d_colclu_n, h->d_colclu, h->ncol, st));
    int last[2];
    HIPCHK(h, hipMemcpyAsync(&last[0], h->d_colclu + h->ncol - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipMemcpyAsync(&last[1], h->d_colclu_n + h->ncol - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    h->nclu = last[0] + last[1];
    if (h->nclu > h->nclu_cap) {
        h->err = "internal: cluster count above bound";
        return MDX_ERR_STATE;
    }
    const int nslot = h->nclu * kClu;
    k_fill_i32<<<cdiv(nslot, 256), 256, 0, st>>>(nslot, h->d_clu_atom, -1);
    k_fill<<<cdiv(n, 256), 256, 0, st>>>(n, h->d_keys2, h->d_vals2, h->d_colstart, h->d_colclu, h->d_clu_atom,
                                          h->d_atom_slot, h->d_clu_col);
    k_cluster_geom<<<cdiv(h->nclu, 256), 256, 0, st>>>(h->nclu, h->d_clu_atom, h->d_x, h->d_y, h->d_z, h->box,
                                                        h->d_anchor, h->d_ext);
    const double rl2 = list_radius2(h);
    const double rl = std::sqrt(rl2);
    const double cwx = h->p.box[0] / h->ncx, cwy = h->p.box[1] / h->ncy;
    const int rx = (int)std::ceil((rl + cwx) / cwx), ry = (int)std::ceil((rl + cwy) / cwy);
    for (int attempt = 0; attempt < 4; ++attempt) {
        k_fill_i32<<<1, 64, 0, st>>>(1, h->d_overflow, 0);
        k_candidates<<<cdiv(h->nclu, 128), 128, 0, st>>>(h->nclu, h->d_anchor, h->d_ext, h->d_clu_col, h->d_colclu,
                                                          h->d_colclu_n, h->ncx, h->ncy, rx, ry, h->box, rl,
                                                          h->cand_cap, h->d_cand, h->d_ncand, h->d_overflow);
        k_masks<<<h->nclu, 64, 0, st>>>(h->nclu, h->d_cand, h->d_ncand, h->cand_cap, h->d_clu_atom, h->d_anchor, h->d_x,
                                        h->d_y, h->d_z, h->d_mol, h->d_moltype, h->d_type_ee, h->d_excl_start,
                                        h->d_excl_atom, h->d_excl_is14, h->box, rl2,
                                        h->p.rshort * h->p.rshort, h->d_list, h->d_nlist, h->list_cap,
                                        h->d_overflow);
        int ovf = 0;
        HIPCHK(h, hipMemcpyAsync(&ovf, h->d_overflow, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        if (ovf == 0) break;
        // grow and retry (the reference stops instead, forces.f:2188-2194)
        const int cap = (ovf + ovf / 4 + 63) / 64 * 64;
        h->cand_cap = std::max(h->cand_cap, cap);
        h->list_cap = h->cand_cap;
        CHK(dalloc(h, &h->d_cand, (size_t)h->nclu_cap * h->cand_cap));
        CHK(dalloc(h, &h->d_list, (size_t)h->nclu_cap * h->list_cap));
        if (attempt == 3) {
            h->err = "pair list overflow";
            return MDX_ERR_LIST_OVERFLOW;
        }
    }
    k_slot_info<<<cdiv(nslot, 256), 256, 0, st>>>(nslot, h->d_clu_atom, h->d_ljtype, h->d_moltype, h->d_type_ee,
                                                   h->d_q, h->d_aux);
    if (h->nspecial > 0)
        k_special_classify<<<cdiv(h->nspecial, 256), 256, 0, st>>>(h->nspecial, h->d_sp_pairs, h->d_sp_kind, h->d_x,
                                                                   h->d_y, h->d_z, h->box, rl2,
                                                                   h->p.rshort * h->p.rshort, h->d_sp_cls);
    if (h->timing) {
        hipEventRecord(h->ev[1], st);
        hipEventSynchronize(h->ev[1]);
        float ms = 0;
        hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
        h->stage_ms[ST_LIST] = ms;
    }
    HIPCHK(h, hipGetLastError());
    h->have_list = true;
    return MDX_OK;
}

mdx_status mdx_list_counts(mdx_engine *h, int64_t *n_short, int64_t *n_long) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) return MDX_ERR_STATE;
    std::vector<int> nl(h->nclu);
    std::vector<PairEntry> L((size_t)h->nclu * h->list_cap);
    std::vector<signed char> cls(std::max(1, h->nspecial));
    HIPCHK(h, hipMemcpy(nl.data(), h->d_nlist, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(L.data(), h->d_list, sizeof(PairEntry) * L.size(), hipMemcpyDeviceToHost));
    if (h->nspecial) HIPCHK(h, hipMemcpy(cls.data(), h->d_sp_cls, h->nspecial, hipMemcpyDeviceToHost));
    int64_t s = 0, l = 0;
    for (int c = 0; c < h->nclu; ++c)
        for (int k = 0; k < nl[c]; ++k) {
            const PairEntry &e = L[(size_t)c * h->list_cap + k];
            s += __builtin_popcountll(e.mask & e.mshort);
            l += __builtin_popcountll(e.mask & ~e.mshort);
        }
    s /= 2;
    l /= 2;
    for (int k = 0; k < h->nspecial; ++k) {
        if (cls[k] == SP_SHORT14 || cls[k] == SP_SHORT) ++s;
        else if (cls[k] == SP_LONG) ++l;
    }
    if (n_short) *n_short = s;
    if (n_long) *n_long = l;
    return MDX_OK;
}

mdx_status mdx_get_list(mdx_engine *h, int32_t *nbs1, int32_t *nbs2, int64_t cap_short, int32_t *nbl1,
                        int32_t *nbl2, int64_t cap_long) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) return MDX_ERR_STATE;
    std::vector<int> nl(h->nclu), ca((size_t)h->nclu * kClu);
    std::vector<PairEntry> L((size_t)h->nclu * h->list_cap);
    std::vector<signed char> cls(std::max(1, h->nspecial));
    HIPCHK(h, hipMemcpy(nl.data(), h->d_nlist, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ca.data(), h->d_clu_atom, sizeof(int) * ca.size(), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(L.data(), h->d_list, sizeof(PairEntry) * L.size(), hipMemcpyDeviceToHost));
    if (h->nspecial) HIPCHK(h, hipMemcpy(cls.data(), h->d_sp_cls, h->nspecial, hipMemcpyDeviceToHost));
    // (key, first, second) with the reference visiting order as key:
    // pair {a > b}: a-b even -> outer a, even strip, step (a-b)/2;
    //               a-b odd  -> outer b, odd strip,  step (a-b+1)/2.
    struct P { uint64_t key; int32_t s1, s2; };
    std::vector<P> S, Lg;
    auto keyof = [](int a, int b) -> uint64_t {
        const int d = a - b;
        const uint64_t outer = (d % 2 == 0) ? a : b;
        const uint64_t strip = (d % 2 == 0) ? 0 : 1;
        const uint64_t step = (d % 2 == 0) ? d / 2 : (d + 1) / 2;
        return (outer << 33) | (strip << 32) | step;
    };
    for (int c = 0; c < h->nclu; ++c)
        for (int k = 0; k < nl[c]; ++k) {
            const PairEntry &e = L[(size_t)c * h->list_cap + k];
            for (int lane = 0; lane < 64; ++lane) {
                if (!((e.mask >> lane) & 1ull)) continue;
                const int ai = ca[c * kClu + (lane >> 3)], aj = ca[e.cj * kClu + (lane & 7)];
                if (ai < aj) continue;
                P q{keyof(ai, aj), ai + 1, aj + 1};
                if ((e.mshort >> lane) & 1ull) S.push_back(q);
                else Lg.push_back(q);
            }
        }
    for (int k = 0; k < h->nspecial; ++k) {
        const int a = h->h_sp_pairs[k].x, b = h->h_sp_pairs[k].y;   // a > b
        const int d = a - b;
        if (cls[k] == SP_SHORT14) {
            // stored (-outer, other) (forces.f:1956-1957, 2127-2128)
            const int outer = (d % 2 == 0) ? a : b, other = (d % 2 == 0) ? b : a;
            S.push_back(P{keyof(a, b), -(outer + 1), other + 1});
        } else if (cls[k] == SP_SHORT) {
            S.push_back(P{keyof(a, b), a + 1, b + 1});
        } else if (cls[k] == SP_LONG) {
            Lg.push_back(P{keyof(a, b), a + 1, b + 1});
        }
    }
    auto lt = [](const P &x, const P &y) { return x.key < y.key; };
    std::sort(S.begin(), S.end(), lt);
    std::sort(Lg.begin(), Lg.end(), lt);
    if ((int64_t)S.size() > cap_short || (int64_t)Lg.size() > cap_long) {
        h->err = "export capacity too small";
        return MDX_ERR_LIST_OVERFLOW;
    }
    for (size_t k = 0; k < S.size(); ++k) { nbs1[k] = S[k].s1; nbs2[k] = S[k].s2; }
    for (size_t k = 0; k < Lg.size(); ++k) { nbl1[k] = Lg[k].s1; nbl2[k] = Lg[k].s2; }
    return MDX_OK;
}

static mdx_status compute_impl(mdx_engine *h, int32_t group, int32_t want, double *dF, double *dG, double *dH,
                               mdx_energies *e) {
    (void)want;
    if (!h->have_list) {
        h->err = "mdx_compute before mdx_build_list";
        return MDX_ERR_STATE;
    }
    if (group < 0 || group > 2) return MDX_ERR_ARG;
    hipStream_t st = h->stream;
    const int n = h->n;
    const int nslot = h->nclu * kClu;
    const bool verlet = h->p.list_mode == MDX_LIST_VERLET;
    const double rc2 = verlet ? h->p.rcut * h->p.rcut : HUGE_VAL;
    // energy partial rows
    const int rows_pairs = h->nclu;
    const int rows_sp = h->nspecial ? cdiv(h->nspecial, 256) : 0;
    const int rows_sk = group != 1 ? cdiv(h->nkv, 256) : 0;
    const int rows_et = (group != 1 && h->neterm) ? cdiv(h->neterm, 256) : 0;
    const int rows_b = (group != 2 && h->nb) ? cdiv(h->nb, 256) : 0;
    const int rows_a = (group != 2 && h->na) ? cdiv(h->na, 256) : 0;
    const int rows_t = (group != 2 && h->nt) ? cdiv(h->nt, 256) : 0;
    const int rows_c = cdiv(n, 256);
    const int total_rows = rows_pairs + rows_sp + rows_sk + rows_et + rows_b + rows_a + rows_t + rows_c;
    CHK(ensure_epart(h, total_rows));
    auto mark = [&](int k) { if (h->timing) hipEventRecord(h->ev[k], st); };
    mark(0);
    k_zero_u64<<<cdiv((size_t)total_rows * Q_N, 256), 256, 0, st>>>((size_t)total_rows * Q_N,
                                                                      (unsigned long long *)h->d_epart);
    k_zero_u64<<<cdiv((size_t)9 * n, 256), 256, 0, st>>>((size_t)9 * n, h->d_fix);
    k_zero_u64<<<cdiv((size_t)6 * n, 256), 256, 0, st>>>((size_t)3 * n, (unsigned long long *)h->d_fs);
    k_zero_u64<<<cdiv((size_t)6 * n, 256), 256, 0, st>>>((size_t)3 * n, (unsigned long long *)h->d_fl);
    k_pack<<<cdiv(nslot, 256), 256, 0, st>>>(nslot, h->d_clu_atom, h->d_anchor, h->d_x, h->d_y, h->d_z, h->d_q,
                                              h->box, h->d_xq);
    if (!h->user_com)
        k_com<<<cdiv(h->nmol, 256), 256, 0, st>>>(h->nmol, h->d_mstart, h->d_mass, h->d_x, h->d_y, h->d_z, h->d_cx,
                                                  h->d_cy, h->d_cz);
    // real-space pairs
    PairArgs pa;
    pa.list = h->d_list;
    pa.nlist = h->d_nlist;
    pa.list_cap = h->list_cap;
    pa.xq = h->d_xq;
    pa.aux = h->d_aux;
    pa.q = h->d_q;
    pa.coulf_f = (float)h->p.coulomb_factor;
    pa.coulf = h->p.coulomb_factor;
    pa.lj = h->d_lj;
    pa.nlj = h->nlj;
    pa.clu_atom = h->d_clu_atom;
    pa.alpha = (float)h->p.alpha;
    pa.rc2 = verlet ? (float)rc2 : INFINITY;
    pa.raex = (float)h->p.ee_raex;
    {
        // Cluster image shifts are minimum images of the anchors; with
        // clusters up to ~5 A across, a pair can need another image only
        // when half the box is within ~10 A of the list radius.
        const double rl = std::sqrt(list_radius2(h));
        const double hmin = 0.5 * std::min(h->p.box[0], std::min(h->p.box[1], h->p.box[2]));
        pa.wrap = hmin < rl + 10.0;
        pa.lx = (float)h->p.box[0]; pa.ly = (float)h->p.box[1]; pa.lz = (float)h->p.box[2];
        pa.ilx = (float)(1.0 / h->p.box[0]); pa.ily = (float)(1.0 / h->p.box[1]); pa.ilz = (float)(1.0 / h->p.box[2]);
    }
    pa.group_mask = 0;
    pa.group = group;
    pa.fs_x = h->d_fs; pa.fs_y = h->d_fs + n; pa.fs_z = h->d_fs + 2 * n;
    pa.fl_x = h->d_fl; pa.fl_y = h->d_fl + n; pa.fl_z = h->d_fl + 2 * n;
    pa.epart = h->d_epart;
    mark(1);
    if (h->p.ee_enabled) k_pairs<true><<<h->nclu, 64, 0, st>>>(pa);
    else k_pairs<false><<<h->nclu, 64, 0, st>>>(pa);
    mark(2);
    unsigned long long *fixG = h->d_fix, *fixHL = h->d_fix + 3 * (size_t)n, *fixHB = h->d_fix + 6 * (size_t)n;
    int row = rows_pairs;
    if (h->nspecial) {
        SpecialArgs sa;
        sa.ns = h->nspecial;
        sa.pairs = h->d_sp_pairs;
        sa.cls = h->d_sp_cls;
        sa.x = h->d_x; sa.y = h->d_y; sa.z = h->d_z;
        sa.atom_moltype = h->d_moltype;
        sa.atom_ljtype = h->d_ljtype;
        sa.q0 = h->d_q0;
        sa.type_ee = h->d_type_ee;
        sa.c14el = h->d_c14el; sa.c14lj = h->d_c14lj;
        sa.a6_0 = h->d_a6_0; sa.b12_0 = h->d_b12_0; sa.a6_14 = h->d_a6_14; sa.b12_14 = h->d_b12_14;
        sa.nlj = h->nlj;
        sa.coulf = h->p.coulomb_factor;
        sa.alpha = h->p.alpha;
        sa.rc2 = rc2;
        sa.box = h->box;
        sa.fscale = h->fscale;
        sa.fixh_x = fixHL; sa.fixh_y = fixHL + n; sa.fixh_z = fixHL + 2 * n;
        sa.fixg_x = fixG; sa.fixg_y = fixG + n; sa.fixg_z = fixG + 2 * n;
        sa.group = group;
        sa.epart = h->d_epart + (size_t)row * Q_N;
        k_special<<<rows_sp, 256, 0, st>>>(sa);
        row += rows_sp;
    }
    mark(3);
    if (group != 1) {
        k_recip_prep<<<cdiv(n, 256), 256, 0, st>>>(n, h->d_x, h->d_y, h->d_z, h->d_q, h->box, h->d_fq,
                                                    h->d_stepz);
        mark(4);
        const dim3 gsk(cdiv(h->nrows, 64), h->nchunk);
        if (h->maxlen <= 16)
            k_sk<16><<<gsk, 64, 0, st>>>(h->nrows, h->d_rows, h->d_rowoff, h->d_wavelen, n, h->chunk, h->d_fq, h->d_stepz,
                                         h->nkv, h->d_part);
        else if (h->maxlen <= 32)
            k_sk<32><<<gsk, 64, 0, st>>>(h->nrows, h->d_rows, h->d_rowoff, h->d_wavelen, n, h->chunk, h->d_fq, h->d_stepz,
                                         h->nkv, h->d_part);
        else if (h->maxlen <= 64)
            k_sk<64><<<gsk, 64, 0, st>>>(h->nrows, h->d_rows, h->d_rowoff, h->d_wavelen, n, h->chunk, h->d_fq, h->d_stepz,
                                         h->nkv, h->d_part);
        else
            k_sk<96><<<gsk, 64, 0, st>>>(h->nrows, h->d_rows, h->d_rowoff, h->d_wavelen, n, h->chunk, h->d_fq, h->d_stepz,
                                         h->nkv, h->d_part);
        mark(5);
        k_sk_reduce<<<rows_sk, 256, 0, st>>>(h->nkv, h->nchunk, h->d_part, h->d_kidx, h->box, h->p.alpha,
                                             h->p.coulomb_factor, h->d_ak, h->d_sk, h->d_epart + (size_t)row * Q_N);
        row += rows_sk;
        mark(6);
        const dim3 gf(cdiv(n, 64), h->nslice);
        k_recip_force<<<gf, 64, 0, st>>>(n, h->nrows, h->rows_per_slice, h->d_rows, h->d_rowoff, h->d_fq, h->d_stepz,
                                         h->d_ak, h->box, h->d_rpart);
        mark(7);
        ScatterArgs sc{h->d_x, h->d_y, h->d_z, h->box, h->fscale, fixG, fixG + n, fixG + 2 * n,
                       h->d_epart + (size_t)row * Q_N};
        if (h->neterm) k_eterms<<<rows_et, 256, 0, st>>>(h->neterm, h->d_et_pairs, h->d_moltype, h->d_q,
                                                         h->p.coulomb_factor, h->p.alpha, sc);
        row += rows_et;
    } else {
        mark(4); mark(5); mark(6); mark(7);
    }
    if (group != 2) {
        ScatterArgs sc{h->d_x, h->d_y, h->d_z, h->box, h->fscale, fixHB, fixHB + n, fixHB + 2 * n, nullptr};
        if (h->nb) {
            sc.epart = h->d_epart + (size_t)row * Q_N;
            k_bonds<<<rows_b, 256, 0, st>>>(h->bonds, sc);
            row += rows_b;
        }
        if (h->na) {
            sc.epart = h->d_epart + (size_t)row * Q_N;
            k_angles<<<rows_a, 256, 0, st>>>(h->angles, sc);
            row += rows_a;
        }
        if (h->nt) {
            sc.epart = h->d_epart + (size_t)row * Q_N;
            k_tors<<<rows_t, 256, 0, st>>>(h->tors, sc);
            row += rows_t;
        }
    }
    mark(8);
    CombineArgs ca;
    ca.n = n;
    ca.group = group;
    ca.fsx = h->d_fs; ca.fsy = h->d_fs + n; ca.fsz = h->d_fs + 2 * n;
    ca.flx = h->d_fl; ca.fly = h->d_fl + n; ca.flz = h->d_fl + 2 * n;
    ca.recip = h->d_rpart;
    ca.nslice = group != 1 ? h->nslice : 0;
    ca.gx = fixG; ca.gy = fixG + n; ca.gz = fixG + 2 * n;
    ca.hlx = fixHL; ca.hly = fixHL + n; ca.hlz = fixHL + 2 * n;
    ca.hbx = fixHB; ca.hby = fixHB + n; ca.hbz = fixHB + 2 * n;
    ca.inv_scale = 1.0 / h->fscale;
    ca.x = h->d_x; ca.y = h->d_y; ca.z = h->d_z;
    ca.atom_mol = h->d_mol;
    ca.cx = h->d_cx; ca.cy = h->d_cy; ca.cz = h->d_cz;
    ca.ox = dF; ca.oy = dF + n; ca.oz = dF + 2 * n;
    ca.Gx = dG; ca.Gy = dG ? dG + n : nullptr; ca.Gz = dG ? dG + 2 * n : nullptr;
    ca.Hx = dH; ca.Hy = dH ? dH + n : nullptr; ca.Hz = dH ? dH + 2 * n : nullptr;
    ca.epart = h->d_epart + (size_t)row * Q_N;
    k_combine<<<rows_c, 256, 0, st>>>(ca);
    row += rows_c;
    k_ereduce<<<Q_N, 256, 0, st>>>(row, h->d_epart, h->d_energy);
    mark(9);
    HIPCHK(h, hipGetLastError());
    if (e) {
        double q[Q_N];
        HIPCHK(h, hipMemcpyAsync(q, h->d_energy, sizeof(q), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        std::memset(e, 0, sizeof(*e));
        e->pe1 = q[Q_PE1]; e->pe2 = q[Q_PE2]; e->pels1 = q[Q_PELS1]; e->pels2 = q[Q_PELS2];
        e->spe = q[Q_SPE] + (group != 1 ? h->spei : 0.0);
        e->pint = q[Q_PINT];
        e->vir1 = q[Q_VIR1]; e->vir2 = q[Q_VIR2]; e->virb = q[Q_VIRB];
        e->virx = q[Q_VIRX]; e->viry = q[Q_VIRY]; e->virz = q[Q_VIRZ];
        e->virfx = q[Q_VIRFX]; e->virfy = q[Q_VIRFY]; e->virfz = q[Q_VIRFZ];
        e->wirs = group != 1 ? q[Q_WIRS] : 0.0;
        e->wirss = group != 2 ? q[Q_WIRSS] : 0.0;
        e->npairs_short = (int64_t)std::llround(q[Q_NPS]);
        e->npairs_long = (int64_t)std::llround(q[Q_NPL]);
        for (int t = 0; t < h->ntypes && group != 1; ++t) e->spee[t] = h->spee_self[t];
    }
    if (h->timing) {
        HIPCHK(h, hipEventSynchronize(h->ev[9]));
        for (int k = ST_PREP; k <= ST_REDUCE; ++k) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, h->ev[k - 1], h->ev[k]);
            h->stage_ms[k] = ms;
        }
    }
    return MDX_OK;
}

mdx_status mdx_compute(mdx_engine *h, int32_t group, int32_t want, double *fx, double *fy, double *fz, double *gx,
                       double *gy, double *gz, double *hx, double *hy, double *hz, mdx_energies *e) {
    if (!h) return MDX_ERR_ARG;
    const int n = h->n;
    double *dF = h->d_out, *dG = h->d_out + 3 * (size_t)n, *dH = h->d_out + 6 * (size_t)n;
    CHK(compute_impl(h, group, want, dF, (gx || gy || gz) ? dG : nullptr, (hx || hy || hz) ? dH : nullptr, e));
    hipStream_t st = h->stream;
    double *outs[9] = {fx, fy, fz, gx, gy, gz, hx, hy, hz};
    for (int k = 0; k < 9; ++k)
        if (outs[k])
            HIPCHK(h, hipMemcpyAsync(outs[k], h->d_out + (size_t)k * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    return MDX_OK;
}

mdx_status mdx_compute_device(mdx_engine *h, int32_t group, int32_t want, double *dfx, double *dfy, double *dfz,
                              mdx_energies *e) {
    if (!h) return MDX_ERR_ARG;
    CHK(compute_impl(h, group, want, h->d_out, nullptr, nullptr, e));
    const size_t b = sizeof(double) * h->n;
    if (dfx) HIPCHK(h, hipMemcpyAsync(dfx, h->d_out, b, hipMemcpyDeviceToDevice, h->stream));
    if (dfy) HIPCHK(h, hipMemcpyAsync(dfy, h->d_out + h->n, b, hipMemcpyDeviceToDevice, h->stream));
    if (dfz) HIPCHK(h, hipMemcpyAsync(dfz, h->d_out + 2 * (size_t)h->n, b, hipMemcpyDeviceToDevice, h->stream));
    if (!e) HIPCHK(h, hipStreamSynchronize(h->stream));
    return MDX_OK;
}

mdx_status mdx_stage_times(mdx_engine *h, double *ms, int32_t n) {
    if (!h || !ms) return MDX_ERR_ARG;
    for (int k = 0; k < n && k < ST_N; ++k) ms[k] = h->stage_ms[k];
    return MDX_OK;
}

}  // extern "C"
