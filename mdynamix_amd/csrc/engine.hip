// Host side of the MI355X force engine: the C ABI of include/mdx.h.
//
// One handle = one system on one HIP stream.  mdx_create copies the
// topology to HBM and precomputes everything static (k-vector rows as in
// INIFUR, forces.f:1356-1449; Ewald self energy; exclusion-correction
// pairs of ETERMS; the special-pair candidates).  mdx_build_list is CHCNB,
// mdx_compute is ALLFORCE / FFORCES / SLFORCES.  All device work of a call
// is queued on the handle's stream; the host waits only where a result
// must come back (list sizes after a rebuild, energies and forces).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "mdx_internal.h"

#include "list.hip"
#include "pair.hip"
#include "recip.hip"
#include "bonded.hip"
#include "md.hip"
#include "furir64.hip"
#include "mdee.hip"

using namespace mdx;

namespace {

// Timed stages (HIP events on the handle's stream).
// Stage k >= 1 lies between events k-1 and k of launch_call.
enum Stage { ST_LIST = 0, ST_PREP, ST_PAIRS, ST_SPECIAL, ST_RPREP, ST_SK, ST_SKRED, ST_RFORCE, ST_BONDED, ST_REDUCE, ST_N };

}  // namespace

struct mdx_batch;

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
    int *d_type_pair = nullptr;
    // per-type energy breakdown rows (allocated on the first detail call)
    double *d_dpart = nullptr;
    int ndet = 0, det_rows = 0;
    double *d_q = nullptr, *d_q0 = nullptr, *d_mass = nullptr, *d_c14el = nullptr, *d_c14lj = nullptr;
    double *d_a6_0 = nullptr, *d_b12_0 = nullptr, *d_a6_14 = nullptr, *d_b12_14 = nullptr;
    float2 *d_lj = nullptr;
    double2 *d_lj64 = nullptr;  // (A6, B12) for the FP64 pair kernel
    int *d_excl_start = nullptr, *d_excl_atom = nullptr;
    signed char *d_excl_is14 = nullptr;
    int *d_mstart = nullptr;
    int *d_comgrp = nullptr;    // molecule groups of the centre-of-mass kernels
    int ncomgrp = 0;
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
    int nb = 0, na = 0, nt = 0, nspc = 0;
    // per-template sums of the bonded terms (mdx_get_terms)
    int ntm = 0, nbtm = 0, natm = 0, nttm = 0;   // templates: all, bond, angle, dihedral
    double2 *d_tv = nullptr;                      // per term (value, energy)
    int *d_tm_start = nullptr, *d_tm_slot = nullptr;
    double *d_tm_w = nullptr, *d_terms = nullptr;
    TermArgs a_terms{};
    std::vector<void *> owned;
    // argument structs of the force kernels, refreshed whenever a list
    // rebuild or box change moves a pointer or count (version counts those)
    PackArgs a_pack{};
    PairArgs a_pair{};
    SpecialArgs a_special{};
    RecipArgs a_recip{};
    EtermArgs a_et{};
    BondArgs a_bonds{};
    SpcArgs a_spc{};
    AngleArgs a_angles{};
    TorsArgs a_tors{};
    long long version = 1;
    int pair_waves = 6;              // wavefronts per SIMD of the plain pair kernel (5 or 6)
    // per-system rows of the energy partials (fixed by capacities)
    int row_pairs = 0, row_pp = 0, row_sp = 0, row_sk = 0, row_et = 0, row_b = 0, row_s = 0, row_a = 0, row_t = 0, row_spr = 0,
        row_c = 0, rows_main = 0, row_bp = 0, row_br = 0, rows_total = 0;
    unsigned long long *d_fixsp = nullptr;        // [3][n] bead springs
    std::map<void *, size_t> dev_sizes;           // live device allocations
    mdx_batch *self = nullptr;                    // batch of one for mdx_compute
    // coordinates
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    bool have_pos = false;
    double *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr;
    bool user_com = false;
    // clusters and lists
    int ncx = 0, ncy = 0, ncol = 0, nclu = 0, nclu_cap = 0;
    double cw_scale = 1.0;           // column width / (8 atoms' volume)^(1/3) (autotuned)
    unsigned long long *d_keys = nullptr, *d_keys2 = nullptr;
    int *d_vals = nullptr, *d_vals2 = nullptr;
    int *d_colcount = nullptr, *d_colstart = nullptr, *d_colclu_n = nullptr, *d_colclu = nullptr;
    int *d_nclu = nullptr;                        // cluster count of the last rebuild (device)
    int *d_clu_atom = nullptr, *d_atom_slot = nullptr, *d_clu_col = nullptr;
    double4 *d_spos = nullptr;                    // per slot: FP64 x, y, z, molecule (list build)
    double *d_anchor = nullptr;
    float *d_ext = nullptr;
    float *d_rbox = nullptr;           // per cluster: lowest, highest stored x, y, z (not wrapped; FP32 outwards)
    int cand_cap = 0, list_cap = 0;
    int *d_cand = nullptr, *d_ncand = nullptr, *d_nlist = nullptr, *d_nshort = nullptr, *d_overflow = nullptr;
    PairEntry *d_list = nullptr;
    // j-reaction records per list slot and the transposed list (k_jgather)
    float *d_rec = nullptr;
    int2 *d_jcnt = nullptr;
    unsigned *d_jstart = nullptr, *d_jdest = nullptr;
    float4 *d_xq = nullptr;
    int2 *d_aux = nullptr;
    bool have_list = false;
    // reciprocal space
    int nkv = 0, nrows = 0, chunk = 512, nchunk = 0, nslice = 1, tiles_per_slice = 0, ntiles = 0;
    int nfrows = 0, npk = 0, mz = 0;
    int tuned_nslice = 0;       // 0: heuristic
    FRow *d_frows = nullptr;
    int4 *d_fpk = nullptr;
    FTile *d_ftiles = nullptr;
    uint16_t *d_wm = nullptr;   // f16 operand of the force pass
    double qabs = 0.0;          // sum |q| (bounds of the force-pass operand)
    float2 *d_fa0 = nullptr;
    double kbox[3] = {0, 0, 0};
    double spei = 0.0;
    std::vector<double> spee_self;
    FRow *d_srows = nullptr;
    float2 *d_tx = nullptr, *d_ty = nullptr, *d_tz = nullptr;
    int kx = 0, ky = 0, tstride = 64;
    double4 *d_ppart = nullptr;
    double2 *d_s0part = nullptr;
    int mzp = 8, nsgroups = 1;
    double4 *d_fq = nullptr;
    double4 *d_stepz = nullptr;

    double *d_rpart = nullptr;
    std::vector<int4> h_kidx;   // (nx, ny, nz) in INIFUR order
    // FP64 precision mode: FURIR workspace (fur_layout of n atoms, nkv k)
    char *d_furws = nullptr;
    // expanded-ensemble energies (mdx_ee_energies), allocated on first use
    double *d_ee_lam = nullptr, *d_ee_rows = nullptr, *d_ee_out = nullptr, *d_ee_kpart = nullptr;
    double4 *d_ee_spart = nullptr;
    int4 *d_ee_kv = nullptr;
    int2 *d_ee_pairs = nullptr;
    int ee_rows_cap = 0, ee_nkv = -1, ee_npairs = -1;
    double ee_kbox[3] = {0, 0, 0};
    double ee_self0 = 0.0;
    // forces
    double *d_fs = nullptr, *d_fl = nullptr;     // [3][n] each
    unsigned long long *d_fix = nullptr;          // [9][n]: G, HL, HB
    unsigned long long *d_fixj = nullptr;         // [slot][8]: pair j reactions HL xyz, G xyz, pad
    // pinned host staging of the host-array copies (positions in, forces
    // out): 9 n doubles, allocated at the first such copy
    double *h_stage = nullptr;
    double *d_out = nullptr;                      // [9][n]: F, G, H
    double *d_epart = nullptr;
    int epart_rows = 0;
    double *d_energy = nullptr;
    double stage_ms[ST_N] = {};
    hipEvent_t ev[2] = {};
    bool timing = true;
    // device-resident MD segment
    bool md_on = false;
    mdx_md_params md{};
    long long mstep = 0;
    int ncons = 0;
    double *d_v = nullptr, *d_o = nullptr, *d_u = nullptr;   // [3][n] each
    double *d_minv = nullptr, *d_mdstate = nullptr, *d_kepart = nullptr;
    int *d_cstart = nullptr;
    int2 *d_cons = nullptr;
    double *d_cons_len2 = nullptr;
    std::vector<int> h_mstart;
    std::vector<double> h_mass;
    // sizing pass of mdx_estimate_bytes: allocations are only counted
    bool dry = false;
    uintptr_t dry_next = 0;
};

// A launch plan over K systems: per kernel an array of argument structs and
// a block map in HBM; one launch per kernel covers all systems.
template <typename T>
struct KPlan {
    std::vector<T> h_args;
    std::vector<int2> h_map;
    T *d_args = nullptr;
    int2 *d_map = nullptr;
    size_t cap_args = 0, cap_map = 0;
    int nblocks() const { return (int)h_map.size(); }
    void clear() { h_args.clear(); h_map.clear(); }
    void add(const T &a, int nblk) {
        const int s = (int)h_args.size();
        h_args.push_back(a);
        for (int b = 0; b < nblk; ++b) h_map.push_back(make_int2(s, b));
    }
};

struct mdx_batch {
    std::vector<mdx_engine *> sys;
    // path-integral MD of a ring batch (mdx_batch_pimd_init)
    bool pimd = false;
    mdx_pimd_params pi{};
    double *d_xi = nullptr;          // [nchain][n] bead thermostat chains
    double **d_vptr = nullptr;       // [4][K] the beads' velocity and x, y, z arrays
    std::vector<long long> seen;     // argument versions last uploaded
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool ring = false;
    double beadfac = 0.0;
    bool use_graph = true;
    bool planned = false;
    int nlj_max = 1;
    bool soft = false;
    int pair_waves = 6;              // the first system's choice
    KPlan<PackArgs> pack, com;
    KPlan<PairArgs> pairs;
    KPlan<PairArgs> pairs64;         // systems in the FP64 precision mode
    std::vector<int> fur64;          // the same systems (FP64 FURIR, one launch set each)
    KPlan<SpecialArgs> special;
    KPlan<RecipArgs> rprep, sk, skred, rforce;
    int mz = 1;
    KPlan<EtermArgs> et;
    KPlan<BondArgs> bonds;
    KPlan<SpcArgs> spc;
    KPlan<AngleArgs> angles;
    KPlan<TorsArgs> tors;
    KPlan<TermArgs> terms;
    KPlan<SpringArgs> springs;
    KPlan<CombineArgs> comb;
    KPlan<ReduceArgs> red, redp;    // energy rows; the pair rows first, on the pair stream
    KPlan<DetailArgs> dred;
    int nmt_max = 1;
    double *d_det = nullptr;         // [K][kMaxDet]
    double *h_det = nullptr;         // pinned copy
    // MD segments
    KPlan<MdArgs> md_kd, md_sh, md_ke, md_th, md_gc, md_tr;
    hipGraph_t md_graph = nullptr;   // one MD step (no list rebuild)
    hipGraphExec_t md_exec = nullptr;
    int md_exec_kind = -1;           // integrator * 65536 + nfreq of md_exec
    double *d_trace = nullptr;
    size_t trace_cap = 0;
    double *d_energy = nullptr;      // [K][Q_N]
    double *h_energy = nullptr;      // pinned copy
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[12] = {};
    // the two side streams of a forked force call and its fork/join events
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
    bool timing = false;
    double stage_ms[ST_N] = {};
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

mdx_status alloc_list(mdx_engine *h);

template <typename T>
mdx_status dalloc(mdx_engine *h, T **p, size_t count) {
    if (h->dry) {
        if (*p) h->dev_sizes.erase((void *)*p);
        if (count == 0) count = 1;
        h->dry_next += ((sizeof(T) * count + 255) / 256 + 1) * 256;
        *p = reinterpret_cast<T *>(h->dry_next);   // never dereferenced
        h->dev_sizes[(void *)*p] = sizeof(T) * count;
        return MDX_OK;
    }
    if (*p) {
        h->dev_sizes.erase((void *)*p);
        hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        h->err = "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed";
        return MDX_ERR_NOMEM;
    }
    h->dev_sizes[(void *)*p] = sizeof(T) * count;
    return MDX_OK;
}

// A device buffer cleared once at allocation.  The accumulators and energy
// rows of a force call are cleared again by the kernels that read them
// (force combine, energy reduction), so every call starts from zeros
// without a fill kernel of its own.
template <typename T>
mdx_status dalloc_zero(mdx_engine *h, T **p, size_t count) {
    CHK(dalloc(h, p, count));
    if (!h->dry) HIPCHK(h, hipMemset(*p, 0, sizeof(T) * std::max<size_t>(count, 1)));
    return MDX_OK;
}

template <typename T>
mdx_status upload(mdx_engine *h, T **p, const T *src, size_t count) {
    *p = nullptr;
    CHK(dalloc(h, p, count));
    if (count && src && !h->dry) HIPCHK(h, hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyHostToDevice));
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

// INIFUR's cube (KMAX, KMAY, KMAZ, forces.f:1370-1372) and cut CUTK2.
void inifur_cube(const mdx_params &p, int &kx, int &ky, int &kz, double &cutk2) {
    cutk2 = 4.0 * (p.alpha * p.alpha) * p.fexp;
    const double cutk = std::sqrt(cutk2);
    const double twopi = 2.0 * kPI;
    kx = (int)(cutk * p.box[0] / twopi + 1.0);
    ky = (int)(cutk * p.box[1] / twopi + 1.0);
    kz = (int)(cutk * p.box[2] / twopi + 1.0);
}

// Size of INIFUR's k set, by the same loops as enumerate_k without
// storing the vectors (no host allocation).
long long count_k(const mdx_params &p) {
    int kx, ky, kz;
    double cutk2;
    inifur_cube(p, kx, ky, kz, cutk2);
    const double twopi = 2.0 * kPI;
    long long n = 0;
    for (int ikz = -kz; ikz <= kz; ++ikz)
        for (int ikx = 0; ikx <= kx; ++ikx) {
            const int ky1 = ikx == 0 ? (ikz > 0 ? 0 : 1) : -ky;
            for (int iky = ky1; iky <= ky; ++iky) {
                const double ax = twopi * ikx / p.box[0], ay = twopi * iky / p.box[1], az = twopi * ikz / p.box[2];
                double rk2 = ax * ax + ay * ay;
                rk2 = rk2 + az * az;
                if (rk2 <= cutk2) ++n;
            }
        }
    return n;
}

bool fur_params_ok(const mdx_params *p) {
    if (!p || !(p->alpha > 0.0) || !(p->fexp > 0.0)) return false;
    for (int k = 0; k < 3; ++k)
        if (!(p->box[k] > 0.0)) return false;
    return true;
}

// Queue the FP64 FURIR kernels on `st` with every buffer in `ws`
// (fur_layout(n, nkv).total bytes).  kbox: the box the k set is enumerated
// at (null: p.box); e4 and qrow: the outputs (either may be null).
void fur_launch(const mdx_params &p, const double *kbox, int n, long long nkv, const double *x, const double *y,
                const double *z, const double *q, double *fx, double *fy, double *fz, double *e4, double *qrow,
                char *ws, hipStream_t st) {
    const FurLayout L = fur_layout(n, (int)nkv);
    FurArgs a{};
    a.n = n;
    a.nkv = (int)nkv;
    a.nchunk = cdiv(std::max(n, 1), kFurChunk);
    mdx_params pk = p;
    if (kbox)
        for (int k = 0; k < 3; ++k) pk.box[k] = kbox[k];
    inifur_cube(pk, a.kmax, a.kmay, a.kmaz, a.cutk2);
    a.klx = pk.box[0]; a.kly = pk.box[1]; a.klz = pk.box[2];
    a.lx = p.box[0]; a.ly = p.box[1]; a.lz = p.box[2];
    a.alp2 = p.alpha * p.alpha;
    a.cfur = 4.0 * kPI * p.coulomb_factor / (p.box[0] * p.box[1] * p.box[2]);
    a.x = x; a.y = y; a.z = z; a.q = q;
    a.fx = fx; a.fy = fy; a.fz = fz; a.e4 = e4; a.qrow = qrow;
    a.kv = reinterpret_cast<int4 *>(ws + L.kv);
    a.spart = reinterpret_cast<double2 *>(ws + L.spart);
    a.ak = reinterpret_cast<double2 *>(ws + L.ak);
    a.epart = reinterpret_cast<double4 *>(ws + L.epart);
    a.count = reinterpret_cast<int *>(ws + L.count);
    const int nkb = cdiv(std::max<long long>(nkv, 1), kFurT);
    k_fur_kset<<<1, 1024, 0, st>>>(a);
    if (nkv > 0 && n > 0) k_fur_sk<<<dim3(nkb, a.nchunk), kFurT, 0, st>>>(a);
    else a.nchunk = 0;
    k_fur_sred<<<nkb, kFurT, 0, st>>>(a);
    if (n > 0) k_fur_force<<<cdiv(n, kFurT), kFurT, 0, st>>>(a);
    k_fur_esum<<<1, kFurT, 0, st>>>(a);
}

// Atom chunks of the structure-factor pass: a multiple of 8 chunks (the
// blocks of one chunk share an XCD) of a multiple of kSkSub atoms, about
// `want` atoms each; the last chunks may be short or empty.
mdx_status set_chunks(mdx_engine *h, int want) {
    h->nchunk = cdiv(cdiv(h->n, std::max(kSkSub, want)), 8) * 8;
    h->chunk = cdiv(cdiv(h->n, h->nchunk), kSkSub) * kSkSub;
    CHK(dalloc(h, &h->d_ppart, (size_t)h->nchunk * std::max(1, h->npk)));
    CHK(dalloc(h, &h->d_s0part, (size_t)h->nchunk * std::max(1, h->nfrows)));
    return MDX_OK;
}

mdx_status setup_kspace(mdx_engine *h) {
    std::vector<int4> ks;
    enumerate_k(h->p, ks);
    // FURIR returns at once for a single molecule (NTYPES = 1 and
    // NSPEC(1) = 1, forces.f:1266): no reciprocal term at all
    if (h->ntypes == 1 && h->nmol == 1) ks.clear();
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
    long long total = 0;
    std::vector<int> off(rows.size());
    for (size_t r = 0; r < rows.size(); ++r) {
        off[r] = (int)total;
        total += rows[r].len;
    }
    if (total != h->nkv) {
        h->err = "k-vector rows are not contiguous";
        return MDX_ERR_NKV;
    }
    h->nrows = (int)rows.size();
    CHK(dalloc(h, &h->d_fq, h->n));
    CHK(dalloc(h, &h->d_stepz, h->n));
    // force pass: rows in (n_x, n_y) order, +-n_z pairs (indices into the
    // row-ordered k arrays above)
    {
        std::vector<int> order(rows.size());
        for (size_t r = 0; r < rows.size(); ++r) order[r] = (int)r;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return rows[a].nx != rows[b].nx ? rows[a].nx < rows[b].nx : rows[a].ny < rows[b].ny;
        });
        std::vector<FRow> fr;
        std::vector<int4> pk;
        int mz = 0;
        for (int r : order) {
            const RecipRow &q = rows[r];
            const int lo = q.nz0, hi = q.nz0 + q.len - 1;
            const int m = std::max(std::abs(lo), std::abs(hi));
            FRow f{};
            f.ix = q.nx;
            f.iy = q.ny;
            f.m = m;
            f.off = (int)pk.size();
            f.k0 = (lo <= 0 && 0 <= hi) ? off[r] - lo : -1;
            // entries padded to a multiple of 4 (zero pairs), so the force
            // kernel reads whole blocks of 4 with unconditional loads
            for (int j = 1; j <= (m + 3) / 4 * 4; ++j) {
                const int kp = (j <= m && lo <= j && j <= hi) ? off[r] + j - lo : -1;
                const int km = (j <= m && lo <= -j && -j <= hi) ? off[r] - j - lo : -1;
                pk.push_back(make_int4(kp, km, j, (int)fr.size()));
            }
            mz = std::max(mz, m);
            fr.push_back(f);
        }
        if (mz > 48) {
            h->err = "k-vector rows with |n_z| above 48 are not supported";
            return MDX_ERR_NKV;
        }
        h->nfrows = (int)fr.size();
        h->npk = (int)pk.size();
        h->mz = mz;
        // phase tables: n_z = 0..mzp-1 (whole N tiles of 8), n_x = 0..kx, |n_y| = 0..ky
        h->mzp = (mz + 1 + 7) / 8 * 8;
        h->kx = h->ky = 0;
        for (const FRow &f : fr) {
            h->kx = std::max(h->kx, std::abs(f.ix));
            h->ky = std::max(h->ky, std::abs(f.iy));
        }
        // structure-factor rows: the same rows sorted by pair count (longest
        // first); pad0 is the row's index in the force-pass order
        std::vector<FRow> sr(fr);
        for (size_t r = 0; r < sr.size(); ++r) sr[r].pad0 = (int)r;
        std::stable_sort(sr.begin(), sr.end(), [](const FRow &a, const FRow &b) { return a.m > b.m; });
        h->nsgroups = cdiv(std::max<size_t>(1, sr.size()), kSkRows);
        CHK(dalloc(h, &h->d_srows, sr.size()));
        if (!h->dry) HIPCHK(h, hipMemcpy(h->d_srows, sr.data(), sizeof(FRow) * sr.size(), hipMemcpyHostToDevice));
        h->tstride = cdiv(h->n, 64) * 64;
        CHK(dalloc(h, &h->d_tx, (size_t)h->tstride * (h->kx + 1)));
        CHK(dalloc(h, &h->d_ty, (size_t)h->tstride * (h->ky + 1)));
        CHK(dalloc(h, &h->d_tz, (size_t)h->tstride * h->mzp));
        // force-pass tiles: up to 4 rows of one n_x with consecutive n_y;
        // the matrix operand of a tile is ceil(longest row / 16) steps of
        // kForceStepH halves (f16 hi and lo of 64 lanes x 8 values), zero
        // where a row has no pair entry, scaled by 2^e so that the bound
        // |j D| <= 4 j AKC(k) sum|q| of its rows stays below 2^13
        const double vol = h->p.box[0] * h->p.box[1] * h->p.box[2];
        const double alp2 = h->p.alpha * h->p.alpha;
        auto row_bound = [&](const FRow &f) {
            double b = 0.0;
            for (int j = 1; j <= f.m; ++j) {
                const double kx = 2.0 * kPI * f.ix / h->p.box[0], ky = 2.0 * kPI * f.iy / h->p.box[1],
                             kz = 2.0 * kPI * j / h->p.box[2];
                const double k2 = kx * kx + ky * ky + kz * kz;
                const double akc = 4.0 * kPI * h->p.coulomb_factor / vol * std::exp(-0.25 * k2 / alp2) / k2;
                b = std::max(b, 4.0 * j * akc * h->qabs);
            }
            return b;
        };
        std::vector<FTile> tl;
        std::vector<double> tbound;
        int woff = 0;
        for (size_t r = 0; r < fr.size(); ++r) {
            FTile *cur = tl.empty() ? nullptr : &tl.back();
            if (!cur || cur->nrow == 4 || fr[r].ix != cur->ix || fr[r].iy != cur->iy0 + cur->nrow) {
                if (cur) woff += cur->nstep * kForceStepH;
                FTile t{};
                t.ix = fr[r].ix;
                t.iy0 = fr[r].iy;
                t.r0 = (int)r;
                tl.push_back(t);
                tbound.push_back(0.0);
                cur = &tl.back();
                cur->woff = woff;
            }
            fr[r].pad1 = cur->woff;
            fr[r].pad2 = cur->nrow++;
            cur->nstep = std::max(cur->nstep, (fr[r].m + 15) / 16);
            tbound.back() = std::max(tbound.back(), row_bound(fr[r]));
        }
        if (!tl.empty()) woff += tl.back().nstep * kForceStepH;
        // scale exponents: per tile (FTile.pad0) and per row (bits 8.. of pad2)
        for (size_t t = 0; t < tl.size(); ++t) {
            int e = 0;
            if (tbound[t] > 0.0) e = 13 - (int)std::ceil(std::log2(tbound[t]));
            e = std::max(-120, std::min(120, e));
            tl[t].pad0 = e;
            for (int q = 0; q < tl[t].nrow; ++q) fr[tl[t].r0 + q].pad2 |= (e + 512) << 8;
        }
        // spare steps after the last region: the force pass reads a whole
        // register set of steps from every region
        woff += 3 * kForceStepH;
        h->ntiles = (int)tl.size();
        CHK(dalloc(h, &h->d_ftiles, std::max<size_t>(1, tl.size())));
        CHK(dalloc(h, &h->d_wm, woff));
        if (!h->dry && !tl.empty())
            HIPCHK(h, hipMemcpy(h->d_ftiles, tl.data(), sizeof(FTile) * tl.size(), hipMemcpyHostToDevice));
        if (!h->dry) HIPCHK(h, hipMemset(h->d_wm, 0, sizeof(uint16_t) * woff));
        CHK(dalloc(h, &h->d_frows, fr.size()));
        CHK(dalloc(h, &h->d_fpk, std::max<size_t>(1, pk.size())));
        CHK(dalloc(h, &h->d_fa0, fr.size() + 4));   // + the force pass's unused reads past the last tile
        if (!h->dry) HIPCHK(h, hipMemcpy(h->d_frows, fr.data(), sizeof(FRow) * fr.size(), hipMemcpyHostToDevice));
        if (!h->dry && !pk.empty())
            HIPCHK(h, hipMemcpy(h->d_fpk, pk.data(), sizeof(int4) * pk.size(), hipMemcpyHostToDevice));
    }
    // about 2,048 structure-factor blocks (row groups x atom chunks) unless tuned
    CHK(set_chunks(h, h->tuned_nslice > 0 ? h->chunk : std::max(256, cdiv(h->n * h->nsgroups, 2048))));
    // force pass: slices of tiles so the grid fills the chip (about 6,000
    // wavefronts of 64 atoms) unless tuned
    const int wav = cdiv(h->n, 16 * kForceTiles);
    h->nslice = h->tuned_nslice > 0 ? h->tuned_nslice : std::max(1, std::min(16, cdiv(6000, std::max(1, wav))));
    h->tiles_per_slice = cdiv(std::max(1, h->ntiles), h->nslice);
    CHK(dalloc(h, &h->d_rpart, (size_t)h->nslice * 3 * h->n));
    if (h->p.precision == MDX_PRECISION_FP64) CHK(dalloc(h, &h->d_furws, fur_layout(h->n, h->nkv).total));
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
    const double cw = std::cbrt(kClu / rho) * h->cw_scale;
    h->ncx = std::max(1, (int)(h->p.box[0] / cw));
    h->ncy = std::max(1, (int)(h->p.box[1] / cw));
    h->ncol = h->ncx * h->ncy;
    h->nclu_cap = h->n / kClu + h->ncol + 1;
    if (h->nclu_cap > (int)kCandIdx) {
        h->err = "too many atoms for the cluster index of the candidate lists";
        return MDX_ERR_ARG;
    }
    CHK(dalloc(h, &h->d_keys, h->n));
    CHK(dalloc(h, &h->d_keys2, h->n));
    CHK(dalloc(h, &h->d_vals, h->n));
    CHK(dalloc(h, &h->d_vals2, h->n));
    CHK(dalloc(h, &h->d_colcount, h->ncol));
    CHK(dalloc(h, &h->d_colstart, h->ncol));
    CHK(dalloc(h, &h->d_colclu_n, h->ncol));
    CHK(dalloc(h, &h->d_colclu, h->ncol));
    CHK(dalloc(h, &h->d_clu_atom, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_spos, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_atom_slot, h->n));
    // the force combine reads atom_slot; valid (zero) before the first list
    if (!h->dry) HIPCHK(h, hipMemset(h->d_atom_slot, 0, sizeof(int) * h->n));
    CHK(dalloc(h, &h->d_clu_col, h->nclu_cap));
    CHK(dalloc(h, &h->d_anchor, (size_t)h->nclu_cap * 3));
    CHK(dalloc(h, &h->d_ext, (size_t)h->nclu_cap * 3));
    CHK(dalloc(h, &h->d_rbox, (size_t)h->nclu_cap * 6));
    CHK(dalloc(h, &h->d_ncand, h->nclu_cap));
    CHK(dalloc(h, &h->d_nlist, h->nclu_cap));
    CHK(dalloc(h, &h->d_nshort, h->nclu_cap));
    CHK(dalloc(h, &h->d_overflow, 1));
    CHK(dalloc(h, &h->d_xq, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_aux, (size_t)h->nclu_cap * kClu));
    CHK(dalloc(h, &h->d_nclu, 1));
    // list capacities from the density: clusters whose boxes come within
    // r_list + 2 cluster extents
    const double rl = h->p.rcut + (h->p.list_mode == MDX_LIST_VERLET ? h->p.skin : 0.0);
    const double reach = rl + 2.0 * cw;
    double est = 4.0 / 3.0 * kPI * reach * reach * reach * rho / kClu * 1.3 + 32;
    est = std::min(est, (double)h->nclu_cap);
    h->cand_cap = ((int)est + 63) / 64 * 64;
    h->list_cap = h->cand_cap;
    CHK(dalloc(h, &h->d_cand, (size_t)h->nclu_cap * h->cand_cap));
    CHK(alloc_list(h));
    CHK(dalloc(h, &h->d_jcnt, h->nclu_cap));
    CHK(dalloc(h, &h->d_jstart, h->nclu_cap));
    // per-call buffers; those sized by the atom count only are kept when
    // the grid changes (the forces F, G, H of the last call stay valid)
    if (!h->d_fs) CHK(dalloc(h, &h->d_fs, (size_t)3 * h->n));
    if (!h->d_fl) CHK(dalloc(h, &h->d_fl, (size_t)3 * h->n));
    if (!h->d_fix) CHK(dalloc_zero(h, &h->d_fix, (size_t)9 * h->n));
    CHK(dalloc_zero(h, &h->d_fixj, (size_t)h->nclu_cap * kClu * 8));
    if (!h->d_out) CHK(dalloc(h, &h->d_out, (size_t)9 * h->n));
    return MDX_OK;
}


// Pair list storage at the current capacity: entries, their j-reaction
// records and the transposed list (list slots are 31-bit indices).
mdx_status alloc_list(mdx_engine *h) {
    const size_t slots = (size_t)h->nclu_cap * h->list_cap;
    if (slots >= (size_t)1 << 31) {
        h->err = "pair list above 2^31 entries";
        return MDX_ERR_LIMIT;
    }
    CHK(dalloc(h, &h->d_list, slots));
    CHK(dalloc(h, &h->d_rec, slots * 2 * kRecUnit));
    CHK(dalloc(h, &h->d_jdest, slots));
    return MDX_OK;
}


// Pinned host staging buffer of the host-array copies (9 n doubles).
mdx_status ensure_stage(mdx_engine *h) {
    if (h->h_stage) return MDX_OK;
    if (hipHostMalloc((void **)&h->h_stage, sizeof(double) * 9 * (size_t)h->n, hipHostMallocDefault) != hipSuccess) {
        h->h_stage = nullptr;
        h->err = "pinned host staging allocation failed";
        return MDX_ERR_NOMEM;
    }
    return MDX_OK;
}


double list_radius2(const mdx_engine *h) {
    const double rl = h->p.rcut + (h->p.list_mode == MDX_LIST_VERLET ? h->p.skin : 0.0);
    return rl * rl;
}


// Rows of the energy partials of one system: fixed by capacities so launch
// shapes do not change at list rebuilds.
void layout_rows(mdx_engine *h) {
    int r = 0;
    h->row_pairs = r; r += h->nclu_cap;
    // the pair rows' sum (formed on the pair kernel's stream), first of
    // the rows the final reduction reads
    h->row_pp = r; r += 1;
    h->row_sp = r; r += h->nspecial ? cdiv(h->nspecial, 256) : 0;
    h->row_sk = r; r += cdiv(h->npk + h->nfrows, kSkRedItems);
    h->row_et = r; r += h->neterm ? cdiv(h->neterm, 256) : 0;
    h->row_b = r; r += h->nb ? cdiv(h->nb, 256) : 0;
    h->row_s = r; r += h->nspc ? cdiv(h->nspc, 256) : 0;
    h->row_a = r; r += h->na ? cdiv(h->na, 256) : 0;
    h->row_t = r; r += h->nt ? cdiv(h->nt, 256) : 0;
    h->row_spr = r; r += cdiv(h->n, 256);
    h->row_c = r; r += cdiv(h->n, 256);
    h->rows_main = r;
    // block sums of the two reductions (pair rows; rows_pp .. rows_main)
    h->row_bp = r; r += cdiv(h->nclu_cap, kRedRows);
    h->row_br = r; r += cdiv(h->rows_main - h->row_pp, kRedRows);
    h->rows_total = r;
}

double *erow(mdx_engine *h, int row) { return h->d_epart + (size_t)row * Q_N; }

void fill_args(mdx_engine *h);

// k-space launch shape: atoms per structure-factor block and row slices of
// the force pass (both change only the work split, not the sums' content
// or order within a k vector and atom).
mdx_status set_kspace_shape(mdx_engine *h, int chunk, int nslice) {
    CHK(set_chunks(h, chunk));
    h->tuned_nslice = std::max(1, std::min(64, nslice));
    h->nslice = h->tuned_nslice;
    h->tiles_per_slice = cdiv(std::max(1, h->ntiles), h->nslice);
    CHK(dalloc(h, &h->d_rpart, (size_t)h->nslice * 3 * h->n));
    fill_args(h);
    return MDX_OK;
}

// Tuned shapes per size class (atoms, k vectors; quarter-octave buckets),
// shared by every handle of the process.
struct TuneKey {
    int n, k;
    bool operator<(const TuneKey &o) const { return n != o.n ? n < o.n : k < o.k; }
};
// A size class's tuned shape: k-space chunk and slices, cluster column
// width scale (0: not tuned)
struct TuneShape {
    int chunk = 0, nslice = 0;
    double cw = 0.0;
    int pw = 0;                      // pair kernel wavefronts per SIMD
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneShape> g_tune;

TuneKey tune_key(const mdx_engine *h) {
    return TuneKey{(int)std::lround(4.0 * std::log2(std::max(1, h->n))),
                   (int)std::lround(4.0 * std::log2(std::max(1, h->nkv)))};
}

void fill_args(mdx_engine *h);

// Rows of the per-type breakdown: one per pair block, then one per
// special block.
mdx_status ensure_detail(mdx_engine *h) {
    const int rows = h->nclu_cap + (h->nspecial ? cdiv(h->nspecial, 256) : 0);
    if (h->d_dpart && rows == h->det_rows) return MDX_OK;
    h->ndet = 2 * h->ntypepairs + 2 * h->ntypes;
    CHK(dalloc(h, &h->d_dpart, (size_t)rows * h->ndet));
    h->det_rows = rows;
    fill_args(h);
    return MDX_OK;
}

// Argument structs of every force kernel from the current state.
void fill_args(mdx_engine *h) {
    const int n = h->n;
    const bool verlet = h->p.list_mode == MDX_LIST_VERLET;
    const double rc2 = verlet ? h->p.rcut * h->p.rcut : HUGE_VAL;
    unsigned long long *fixG = h->d_fix, *fixHL = h->d_fix + 3 * (size_t)n, *fixHB = h->d_fix + 6 * (size_t)n;
    PackArgs &pk = h->a_pack;
    pk.nclu = h->have_list ? h->nclu : 0;
    pk.clu_atom = h->d_clu_atom;
    pk.anchor = h->d_anchor;
    pk.x = h->d_x; pk.y = h->d_y; pk.z = h->d_z; pk.q = h->d_q;
    pk.box = h->box;
    pk.xq = h->d_xq;
    pk.nmol = h->nmol;
    pk.user_com = h->user_com ? 1 : 0;
    pk.mstart = h->d_mstart;
    pk.comgrp = h->d_comgrp;
    pk.mass = h->d_mass;
    pk.cx = h->d_cx; pk.cy = h->d_cy; pk.cz = h->d_cz;

    PairArgs &pa = h->a_pair;
    pa.nclu = h->have_list ? h->nclu : 0;
    pa.list = h->d_list;
    pa.nlist = h->d_nlist;
    pa.nshort = h->d_nshort;
    pa.list_cap = h->list_cap;
    pa.xq = h->d_xq;
    pa.aux = h->d_aux;
    pa.q = h->d_q;
    pa.coulf_f = (float)h->p.coulomb_factor;
    pa.coulf = h->p.coulomb_factor;
    pa.lj = h->d_lj;
    pa.lj64 = h->d_lj64;
    pa.nlj = h->nlj;
    pa.clu_atom = h->d_clu_atom;
    pa.alpha = (float)h->p.alpha;
    pa.rc2 = verlet ? (float)rc2 : INFINITY;
    {
        // FP32 r^2 from anchor-relative coordinates and an FP32 image shift
        // (|shift| <= L/2): |error| <= about 2 r_c L 2^-24 + 3 r_c^2 2^-24;
        // the band is eight times that
        const double lmax = std::max(h->p.box[0], std::max(h->p.box[1], h->p.box[2]));
        pa.band = verlet ? (float)(1e-6 * h->p.rcut * (lmax + 3.0 * h->p.rcut)) : 0.f;
        pa.rlo = pa.band > 0.f ? std::nextafter((float)(rc2 - pa.band), 0.f) : pa.rc2;
        pa.rhi = pa.band > 0.f ? std::nextafter((float)(rc2 + pa.band), INFINITY) : pa.rc2;
        pa.rc2d = rc2;
        pa.x = h->d_x; pa.y = h->d_y; pa.z = h->d_z;
        pa.box = h->box;
    }
    pa.raex = (float)h->p.ee_raex;
    pa.alphad = h->p.alpha;
    pa.raexd = h->p.ee_raex;
    pa.soft = h->p.ee_enabled ? 1 : 0;
    {
        pa.lx = (float)h->p.box[0]; pa.ly = (float)h->p.box[1]; pa.lz = (float)h->p.box[2];
        pa.ilx = (float)(1.0 / h->p.box[0]); pa.ily = (float)(1.0 / h->p.box[1]); pa.ilz = (float)(1.0 / h->p.box[2]);
    }
    pa.fs_x = h->d_fs; pa.fs_y = h->d_fs + n; pa.fs_z = h->d_fs + 2 * n;
    pa.fl_x = h->d_fl; pa.fl_y = h->d_fl + n; pa.fl_z = h->d_fl + 2 * n;
    pa.rec = h->d_rec;
    pa.jstart = h->d_jstart;
    pa.jcnt = h->d_jcnt;
    pa.jdest = h->d_jdest;
    pa.jf = h->d_fixj;
    pa.fscale = h->fscale;
    pa.epart = erow(h, h->row_pairs);
    pa.atom_mol = h->d_mol;
    pa.atom_moltype = h->d_moltype;
    pa.type_pair = h->d_type_pair;
    pa.nmt = h->ntypes;
    pa.ntp = h->ntypepairs;
    pa.ndet = h->ndet;
    pa.dpart = h->d_dpart;

    SpecialArgs &sa = h->a_special;
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
    sa.epart = erow(h, h->row_sp);
    sa.nmt = h->ntypes;
    sa.ntp = h->ntypepairs;
    sa.ndet = h->ndet;
    sa.dpart = h->d_dpart ? h->d_dpart + (size_t)h->nclu_cap * h->ndet : nullptr;

    RecipArgs &ra = h->a_recip;
    ra.natoms = n;
    ra.x = h->d_x; ra.y = h->d_y; ra.z = h->d_z; ra.q = h->d_q;
    ra.box = h->box;
    ra.fq = h->d_fq;
    ra.stepz = h->d_stepz;
    ra.tx = h->d_tx;
    ra.ty = h->d_ty;
    ra.tz = h->d_tz;
    ra.kx = h->kx;
    ra.tstride = h->tstride;
    ra.ky = h->ky;
    ra.mzp = h->mzp;
    ra.chunk = h->chunk;
    ra.nchunk = h->nchunk;
    ra.nkv = h->nkv;
    ra.alpha = h->p.alpha;
    ra.coulf = h->p.coulomb_factor;
    ra.nsgroups = h->nsgroups;
    ra.srows = h->d_srows;
    ra.ppart = h->d_ppart;
    ra.s0part = h->d_s0part;
    ra.nslice = h->nslice;
    ra.tiles_per_slice = h->tiles_per_slice;
    ra.nfrows = h->nfrows;
    ra.npk = h->npk;
    ra.frows = h->d_frows;
    ra.fpk = h->d_fpk;
    ra.ntiles = h->ntiles;
    ra.ftiles = h->d_ftiles;
    ra.wm = h->d_wm;
    ra.fa0 = h->d_fa0;
    ra.natom_blocks = cdiv(n, kForceBlock);
    ra.rpart = h->d_rpart;
    ra.epart = erow(h, h->row_sk);

    ScatterArgs slow{h->d_x, h->d_y, h->d_z, h->box, h->fscale, fixG, fixG + n, fixG + 2 * n, erow(h, h->row_et)};
    h->a_et.np = h->neterm;
    h->a_et.pairs = h->d_et_pairs;
    h->a_et.q = h->d_q;
    h->a_et.coulf = h->p.coulomb_factor;
    h->a_et.alpha = h->p.alpha;
    h->a_et.s = slow;
    ScatterArgs fast{h->d_x, h->d_y, h->d_z, h->box, h->fscale, fixHB, fixHB + n, fixHB + 2 * n, nullptr};
    h->a_bonds.s = fast;
    h->a_bonds.s.epart = erow(h, h->row_b);
    h->a_spc.s = fast;
    h->a_spc.s.epart = erow(h, h->row_s);
    h->a_angles.s = fast;
    h->a_angles.s.epart = erow(h, h->row_a);
    h->a_tors.s = fast;
    h->a_tors.s.epart = erow(h, h->row_t);
    h->version++;
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

static mdx_status create_impl(const mdx_topology *t, const mdx_params *p, mdx_engine **out, bool dry) {
    if (!t || !p || !out) return MDX_ERR_ARG;
    *out = nullptr;
    mdx_engine *h = new mdx_engine();
    h->dry = dry;
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
    if (p->precision != MDX_PRECISION_MIXED && p->precision != MDX_PRECISION_FP64) {
        h->err = "precision must be MDX_PRECISION_MIXED or MDX_PRECISION_FP64";
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
        if (t->nspc < 0) return bad("nspc");
        for (int k = 0; k < 3 * t->nspc; ++k)
            if (t->spc_atom[k] < 0 || t->spc_atom[k] >= n) return bad("spc_atom");
        for (int k = 0; k < t->nspc; ++k)
            if (t->spc_par[12 * k + 10] != 1.0 && t->spc_par[12 * k + 10] != 2.0) return bad("spc_par IPOT");
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
    if (!dry) {
        if (hipStreamCreate(&h->stream) != hipSuccess) {
            h->err = "hipStreamCreate failed";
            return fail(MDX_ERR_HIP);
        }
        for (int k = 0; k < 2; ++k) hipEventCreate(&h->ev[k]);
    }
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
        (s = upload(h, &h->d_type_ee, t->type_ee, h->ntypes)) ||
        (s = upload(h, &h->d_type_pair, t->type_pair, h->ntypes * h->ntypes)) || (s = upload(h, &h->d_q, t->atom_charge, n)) ||
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
    // LJ only where B12 != 0, as LOCALF/FORCES test it (forces.f:1160)
    std::vector<float2> lj((size_t)h->nlj * h->nlj);
    std::vector<double2> lj64(lj.size());
    for (size_t k = 0; k < lj.size(); ++k) {
        const double a6 = t->b12[k] != 0.0 ? t->a6[k] : 0.0;
        lj[k] = make_float2((float)(6.0 * a6), (float)(12.0 * t->b12[k]));
        lj64[k] = make_double2(a6, t->b12[k]);
    }
    if ((s = upload(h, &h->d_lj, lj.data(), lj.size()))) return fail(s);
    if ((s = upload(h, &h->d_lj64, lj64.data(), lj64.size()))) return fail(s);
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
    h->h_mstart = mstart;
    // centre-of-mass groups: consecutive molecules, at most 64 of them and
    // at most kComAtoms atoms (a larger molecule alone)
    {
        std::vector<int> g(1, 0);
        for (int m = 0; m < h->nmol; ++m) {
            const int first = g.back();
            if (m > first && (m - first == 64 || mstart[m + 1] - mstart[first] > kComAtoms)) g.push_back(m);
        }
        g.push_back(h->nmol);
        h->ncomgrp = (int)g.size() - 1;
        if ((s = upload(h, &h->d_comgrp, g.data(), g.size()))) return fail(s);
    }
    h->h_mass = mass;
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
    h->a_bonds.n = h->nb;
    if ((s = upload(h, &ip, t->bond_i, h->nb))) return fail(s); h->a_bonds.bi = ip;
    if ((s = upload(h, &ip, t->bond_j, h->nb))) return fail(s); h->a_bonds.bj = ip;
    if ((s = upload(h, &ip, t->bond_param, h->nb))) return fail(s); h->a_bonds.bpar = ip;
    if ((s = upload(h, &ip, t->bp_morse, t->nbondparams))) return fail(s); h->a_bonds.morse = ip;
    if ((s = upload(h, &dp, t->bp_r0, t->nbondparams))) return fail(s); h->a_bonds.r0 = dp;
    if ((s = upload(h, &dp, t->bp_k, t->nbondparams))) return fail(s); h->a_bonds.k = dp;
    if ((s = upload(h, &dp, t->bp_d, t->nbondparams))) return fail(s); h->a_bonds.d = dp;
    if ((s = upload(h, &dp, t->bp_rho, t->nbondparams))) return fail(s); h->a_bonds.rho = dp;
    h->nspc = t->nspc;
    h->a_spc.n = h->nspc;
    if ((s = upload(h, &ip, t->spc_atom, 3 * (size_t)h->nspc))) return fail(s); h->a_spc.atom = ip;
    if ((s = upload(h, &dp, t->spc_par, 12 * (size_t)h->nspc))) return fail(s); h->a_spc.par = dp;
    h->a_angles.n = h->na;
    if ((s = upload(h, &ip, t->angle_i, h->na))) return fail(s); h->a_angles.ai = ip;
    if ((s = upload(h, &ip, t->angle_j, h->na))) return fail(s); h->a_angles.aj = ip;
    if ((s = upload(h, &ip, t->angle_k, h->na))) return fail(s); h->a_angles.ak = ip;
    if ((s = upload(h, &ip, t->angle_param, h->na))) return fail(s); h->a_angles.apar = ip;
    if ((s = upload(h, &dp, t->ap_theta0, t->nangleparams))) return fail(s); h->a_angles.theta0 = dp;
    if ((s = upload(h, &dp, t->ap_k, t->nangleparams))) return fail(s); h->a_angles.k = dp;
    h->a_tors.n = h->nt;
    if ((s = upload(h, &ip, t->tors_i, h->nt))) return fail(s); h->a_tors.ti = ip;
    if ((s = upload(h, &ip, t->tors_j, h->nt))) return fail(s); h->a_tors.tj = ip;
    if ((s = upload(h, &ip, t->tors_k, h->nt))) return fail(s); h->a_tors.tk = ip;
    if ((s = upload(h, &ip, t->tors_l, h->nt))) return fail(s); h->a_tors.tl = ip;
    if ((s = upload(h, &ip, t->tors_param, h->nt))) return fail(s); h->a_tors.tpar = ip;
    if ((s = upload(h, &ip, t->tp_kind, t->ntorsparams))) return fail(s); h->a_tors.kind = ip;
    if ((s = upload(h, &ip, t->tp_mult, t->ntorsparams))) return fail(s); h->a_tors.mult = ip;
    if ((s = upload(h, &dp, t->tp_k, t->ntorsparams))) return fail(s); h->a_tors.k = dp;
    if ((s = upload(h, &dp, t->tp_delta, t->ntorsparams))) return fail(s); h->a_tors.delta = dp;
    if ((s = upload(h, &dp, t->tp_k1, t->ntorsparams))) return fail(s); h->a_tors.k1 = dp;
    if ((s = upload(h, &dp, t->tp_k2, t->ntorsparams))) return fail(s); h->a_tors.k2 = dp;
    if ((s = upload(h, &dp, t->tp_k3, t->ntorsparams))) return fail(s); h->a_tors.k3 = dp;
    if ((s = upload(h, &dp, t->tp_k4, t->ntorsparams))) return fail(s); h->a_tors.k4 = dp;
    if ((s = upload(h, &dp, t->tp_k5, t->ntorsparams))) return fail(s); h->a_tors.k5 = dp;
    // per-template sums (BR/BE, AR/AE, TR/TE): term slots are the bonds,
    // the STRBND records' three bonds, the angles, the dihedrals; templates
    // the bond, angle and dihedral parameter sets in that order; each term
    // counts 1 / NSPEC of its molecule type (forces.f:140, 240, 345, 560)
    {
        std::vector<int> nspec(h->ntypes, 0);
        for (int i = 0; i < n; ++i)
            if (i == 0 || t->atom_mol[i] != t->atom_mol[i - 1]) ++nspec[t->atom_moltype[i]];
        h->nbtm = t->nbondparams;
        h->natm = t->nangleparams;
        h->nttm = t->ntorsparams;
        h->ntm = h->nbtm + h->natm + h->nttm;
        const int nslot = h->nb + 3 * h->nspc + h->na + h->nt;
        std::vector<int> tmpl(nslot, -1);
        std::vector<double> w(nslot, 0.0);
        auto put = [&](int slot, int tm, int atom) {
            tmpl[slot] = tm;
            w[slot] = 1.0 / (double)nspec[t->atom_moltype[atom]];
        };
        for (int k = 0; k < h->nb; ++k) put(k, t->bond_param[k], t->bond_i[k]);
        for (int r = 0; r < h->nspc && t->spc_bond; ++r) {
            if (t->spc_bond[r] < 0 || t->spc_bond[r] + 2 >= h->nbtm) {
                h->err = "index out of range: spc_bond";
                return fail(MDX_ERR_ARG);
            }
            for (int q = 0; q < 3; ++q) put(h->nb + 3 * r + q, t->spc_bond[r] + q, t->spc_atom[3 * r]);
        }
        const int a0 = h->nb + 3 * h->nspc, t0 = a0 + h->na;
        for (int k = 0; k < h->na; ++k) put(a0 + k, h->nbtm + t->angle_param[k], t->angle_i[k]);
        for (int k = 0; k < h->nt; ++k) put(t0 + k, h->nbtm + h->natm + t->tors_param[k], t->tors_i[k]);
        std::vector<int> start(h->ntm + 1, 0), slot;
        for (int k = 0; k < nslot; ++k)
            if (tmpl[k] >= 0) ++start[tmpl[k] + 1];
        for (int m = 0; m < h->ntm; ++m) start[m + 1] += start[m];
        slot.assign(start[h->ntm], 0);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int k = 0; k < nslot; ++k)
            if (tmpl[k] >= 0) slot[fill[tmpl[k]]++] = k;
        const std::vector<double> zeros(2 * (size_t)std::max(1, h->ntm), 0.0);
        if ((s = dalloc(h, &h->d_tv, std::max(1, nslot))) ||
            (s = upload(h, &h->d_tm_start, start.data(), start.size())) ||
            (s = upload(h, &h->d_tm_slot, slot.data(), slot.size())) ||
            (s = upload(h, &h->d_tm_w, w.data(), w.size())) ||
            (s = upload(h, &h->d_terms, zeros.data(), zeros.size())))
            return fail(s);
        h->a_terms = TermArgs{h->ntm, h->d_tm_start, h->d_tm_slot, h->d_tm_w, h->d_tv, h->d_terms};
        h->a_bonds.tv = h->d_tv;
        h->a_spc.tv = h->d_tv + h->nb;
        h->a_angles.tv = h->d_tv + a0;
        h->a_tors.tv = h->d_tv + t0;
    }
    // coordinates, COMs, energies
    if ((s = dalloc(h, &h->d_x, n)) || (s = dalloc(h, &h->d_y, n)) || (s = dalloc(h, &h->d_z, n)) ||
        (s = dalloc(h, &h->d_cx, h->nmol)) || (s = dalloc(h, &h->d_cy, h->nmol)) ||
        (s = dalloc(h, &h->d_cz, h->nmol)) || (s = dalloc(h, &h->d_energy, Q_N)))
        return fail(s);
    self_energy(h, t);
    h->qabs = 0.0;
    for (int i = 0; i < n; ++i) h->qabs += std::fabs(t->atom_charge[i]);
    if ((s = setup_kspace(h))) return fail(s);
    TuneShape shape;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(tune_key(h));
        if (it != g_tune.end()) shape = it->second;
    }
    if (shape.cw > 0.0) h->cw_scale = shape.cw;
    if (shape.pw > 0) h->pair_waves = shape.pw;
    if ((s = alloc_clusters(h))) return fail(s);
    if ((s = dalloc_zero(h, &h->d_fixsp, (size_t)3 * n))) return fail(s);
    layout_rows(h);
    if ((s = dalloc_zero(h, &h->d_epart, (size_t)h->rows_total * Q_N))) return fail(s);
    fill_args(h);
    if (dry) {
        *out = h;
        return MDX_OK;
    }
    if (shape.chunk > 0 && (s = set_kspace_shape(h, shape.chunk, shape.nslice))) return fail(s);
    mdx_batch *self = nullptr;
    mdx_engine *one[1] = {h};
    if ((s = mdx_batch_create(one, 1, &self))) return fail(s);
    hipStreamDestroy(self->stream);
    self->stream = h->stream;
    self->own_stream = false;
    self->use_graph = false;
    self->timing = true;
    h->self = self;
    *out = h;
    return MDX_OK;
}

mdx_status mdx_create(const mdx_topology *t, const mdx_params *p, mdx_engine **out) {
    return create_impl(t, p, out, false);
}

mdx_status mdx_estimate_bytes(const mdx_topology *t, const mdx_params *p, int32_t with_md, int64_t *bytes) {
    if (!bytes) return MDX_ERR_ARG;
    mdx_engine *h = nullptr;
    const mdx_status s = create_impl(t, p, &h, true);
    int64_t total = 0;
    if (h) {
        for (auto &kv : h->dev_sizes) total += (int64_t)kv.second;
        // batch of one (energies, breakdown) and MD segment state
        total += (int64_t)sizeof(double) * (Q_N + kMaxDet);
        if (with_md) total += (int64_t)sizeof(double) * (10LL * h->n + MD_NSTATE + cdiv(h->n, 256) + h->nmol);
        delete h;
    }
    *bytes = total;
    return s;
}

mdx_status mdx_destroy(mdx_engine *h) {
    if (!h) return MDX_OK;
    if (h->dry) {
        delete h;
        return MDX_OK;
    }
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->self) mdx_batch_destroy(h->self);
    for (auto &kv : h->dev_sizes) hipFree(kv.first);
    if (h->h_stage) hipHostFree(h->h_stage);
    for (int k = 0; k < 2; ++k)
        if (h->ev[k]) hipEventDestroy(h->ev[k]);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return MDX_OK;
}

mdx_status mdx_set_positions(mdx_engine *h, const double *x, const double *y, const double *z) {
    if (!h || !x || !y || !z) return MDX_ERR_ARG;
    const size_t n = h->n, b = sizeof(double) * n;
    CHK(ensure_stage(h));
    // through pinned memory: pageable copies are staged by the runtime in
    // small synchronous pieces
    HIPCHK(h, hipStreamSynchronize(h->stream));   // the stage may still feed a copy
    std::memcpy(h->h_stage, x, b);
    std::memcpy(h->h_stage + n, y, b);
    std::memcpy(h->h_stage + 2 * n, z, b);
    HIPCHK(h, hipMemcpyAsync(h->d_x, h->h_stage, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_y, h->h_stage + n, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_z, h->h_stage + 2 * n, b, hipMemcpyHostToDevice, h->stream));
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
        fill_args(h);
        return MDX_OK;
    }
    const size_t b = sizeof(double) * h->nmol;
    HIPCHK(h, hipMemcpyAsync(h->d_cx, x, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_cy, y, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_cz, z, b, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->user_com = true;
    fill_args(h);
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
    if (redo) {
        CHK(setup_kspace(h));
        layout_rows(h);
        CHK(dalloc_zero(h, &h->d_epart, (size_t)h->rows_total * Q_N));
    }
    h->have_list = false;
    fill_args(h);
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
    // atoms binned into x-y columns and sorted by z inside each (the
    // column's (z code, atom index) order); clusters of 8 per column.  The
    // cluster count stays on the device: every later kernel reads it and
    // the grids are sized by the capacity n / 8 + ncol + 1, which bounds
    // the sum over columns of ceil(count / 8)
    k_fill_i32<<<cdiv(h->ncol, 256), 256, 0, st>>>(h->ncol, h->d_colcount, 0);
    // per-block column counts in LDS up to 16,384 columns (64 KiB)
    const int lds_cols = h->ncol <= 16384 ? h->ncol : 0;
    k_bin<<<cdiv(n, 256), 256, sizeof(int) * lds_cols, st>>>(n, h->d_x, h->d_y, h->d_z, h->box, h->ncx, h->ncy,
                                                             lds_cols, h->d_keys, h->d_vals, h->d_colcount);
    k_scan_int<<<1, 1024, 0, st>>>(h->ncol, h->d_colcount, h->d_colstart, nullptr);
    k_scatter<<<cdiv(n, 256), 256, 0, st>>>(n, h->d_keys, h->d_vals, h->d_colstart, h->d_keys2);
    k_col_sort<<<h->ncol, 256, 0, st>>>(h->d_colcount, h->d_colstart, h->d_keys2, h->d_keys, h->d_vals2);
    k_col_clusters<<<cdiv(h->ncol, 256), 256, 0, st>>>(h->ncol, h->d_colcount, h->d_colclu_n);
    k_scan_int<<<1, 1024, 0, st>>>(h->ncol, h->d_colclu_n, h->d_colclu, h->d_nclu);
    const int cap = h->nclu_cap, nslot_cap = cap * kClu;
    k_fill_i32<<<cdiv(nslot_cap, 256), 256, 0, st>>>(nslot_cap, h->d_clu_atom, -1);
    k_fill<<<cdiv(n, 256), 256, 0, st>>>(n, h->d_keys, h->d_vals2, h->d_colstart, h->d_colclu, h->d_clu_atom,
                                          h->d_atom_slot, h->d_clu_col);
    k_cluster_geom<<<cdiv(cap, 256), 256, 0, st>>>(h->d_nclu, h->d_clu_atom, h->d_x, h->d_y, h->d_z, h->box,
                                                    h->d_anchor, h->d_ext, h->d_rbox);
    k_slot_pos<<<cdiv(nslot_cap, 256), 256, 0, st>>>(h->d_nclu, h->d_clu_atom, h->d_x, h->d_y, h->d_z, h->d_mol,
                                                      h->d_spos);
    const double rl2 = list_radius2(h);
    const double rl = std::sqrt(rl2);
    const double cwx = h->p.box[0] / h->ncx, cwy = h->p.box[1] / h->ncy;
    const int rx = (int)std::ceil((rl + cwx) / cwx), ry = (int)std::ceil((rl + cwy) / cwy);
    for (int attempt = 0; attempt < 4; ++attempt) {
        k_fill_i32<<<1, 64, 0, st>>>(1, h->d_overflow, 0);
        k_candidates<<<cap, 64, 0, st>>>(h->d_nclu, h->d_anchor, h->d_ext, h->d_rbox, h->d_clu_col, h->d_colclu,
                                                          h->d_colclu_n, h->ncx, h->ncy, rx, ry, h->box, rl,
                                                          h->cand_cap, h->d_cand, h->d_ncand, h->d_overflow);
        k_masks<<<cap, 64, 0, st>>>(h->d_nclu, h->d_cand, h->d_ncand, h->cand_cap, h->d_clu_atom, h->d_spos,
                                        h->d_anchor, h->d_ext, rl, h->d_moltype, h->d_type_ee, h->d_excl_start,
                                        h->d_excl_atom, h->d_excl_is14, h->box, rl2,
                                        h->p.rshort * h->p.rshort, h->d_list, h->d_nlist, h->d_nshort,
                                        h->list_cap, h->d_overflow);
        int got[2] = {0, 0};   // overflow, cluster count: the rebuild's one host synchronisation
        HIPCHK(h, hipMemcpyAsync(&got[0], h->d_overflow, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipMemcpyAsync(&got[1], h->d_nclu, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        h->nclu = got[1];
        if (h->nclu > h->nclu_cap) {
            h->err = "internal: cluster count above bound";
            return MDX_ERR_STATE;
        }
        const int ovf = got[0];
        if (ovf == 0) break;
        // grow and retry (the reference stops instead, forces.f:2188-2194)
        const int cap = (ovf + ovf / 4 + 63) / 64 * 64;
        h->cand_cap = std::max(h->cand_cap, cap);
        h->list_cap = h->cand_cap;
        CHK(dalloc(h, &h->d_cand, (size_t)h->nclu_cap * h->cand_cap));
        CHK(alloc_list(h));
        if (attempt == 3) {
            h->err = "pair list overflow";
            return MDX_ERR_LIST_OVERFLOW;
        }
    }
    // transposed list for the j-reaction sums
    k_fill_i32<<<cdiv(2 * h->nclu, 256), 256, 0, st>>>(2 * h->nclu, (int *)h->d_jcnt, 0);
    k_jcount<<<h->nclu, 64, 0, st>>>(h->d_nclu, h->d_list, h->d_nlist, h->d_nshort, h->list_cap, h->d_jcnt,
                                     h->d_jdest);
    k_scan_excl<<<1, 1024, 0, st>>>(h->d_nclu, h->d_jcnt, h->d_jstart);
    k_jplace<<<h->nclu, 64, 0, st>>>(h->d_nclu, h->d_list, h->d_nlist, h->d_nshort, h->list_cap, h->d_jstart,
                                     h->d_jcnt, h->d_jdest);
    k_slot_info<<<cdiv(nslot_cap, 256), 256, 0, st>>>(h->d_nclu, h->d_clu_atom, h->d_ljtype, h->d_moltype, h->d_type_ee,
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
    fill_args(h);
    return MDX_OK;
}

// Entry k of an i-cluster's list: the ns entries with a short-class pair
// from the front, the rest from the back (k_masks).
inline size_t list_slot(int k, int ns, int cap) { return k < ns ? (size_t)k : (size_t)(cap - 1 - (k - ns)); }

mdx_status mdx_list_counts(mdx_engine *h, int64_t *n_short, int64_t *n_long) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) return MDX_ERR_STATE;
    std::vector<int> nl(h->nclu), ns(h->nclu);
    std::vector<PairEntry> L((size_t)h->nclu * h->list_cap);
    std::vector<signed char> cls(std::max(1, h->nspecial));
    HIPCHK(h, hipMemcpy(nl.data(), h->d_nlist, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ns.data(), h->d_nshort, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(L.data(), h->d_list, sizeof(PairEntry) * L.size(), hipMemcpyDeviceToHost));
    if (h->nspecial) HIPCHK(h, hipMemcpy(cls.data(), h->d_sp_cls, h->nspecial, hipMemcpyDeviceToHost));
    int64_t s = 0, l = 0;
    for (int c = 0; c < h->nclu; ++c)
        for (int k = 0; k < nl[c]; ++k) {
            const PairEntry &e = L[(size_t)c * h->list_cap + list_slot(k, ns[c], h->list_cap)];
            s += __builtin_popcountll(e.mask & e.mshort);
            l += __builtin_popcountll(e.mask & ~e.mshort);
        }
    for (int k = 0; k < h->nspecial; ++k) {
        if (cls[k] == SP_SHORT14 || cls[k] == SP_SHORT) ++s;
        else if (cls[k] == SP_LONG) ++l;
    }
    if (n_short) *n_short = s;
    if (n_long) *n_long = l;
    return MDX_OK;
}

mdx_status mdx_list_entries(mdx_engine *h, int64_t *entries, int64_t *front_entries) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) return MDX_ERR_STATE;
    std::vector<int> nl(h->nclu), ns(h->nclu);
    HIPCHK(h, hipMemcpy(nl.data(), h->d_nlist, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ns.data(), h->d_nshort, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    int64_t e = 0, f = 0;
    for (int c = 0; c < h->nclu; ++c) { e += nl[c]; f += ns[c]; }
    if (entries) *entries = e;
    if (front_entries) *front_entries = f;
    return MDX_OK;
}

mdx_status mdx_get_list(mdx_engine *h, int32_t *nbs1, int32_t *nbs2, int64_t cap_short, int32_t *nbl1,
                        int32_t *nbl2, int64_t cap_long) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) return MDX_ERR_STATE;
    std::vector<int> nl(h->nclu), ns(h->nclu), ca((size_t)h->nclu * kClu);
    std::vector<PairEntry> L((size_t)h->nclu * h->list_cap);
    std::vector<signed char> cls(std::max(1, h->nspecial));
    HIPCHK(h, hipMemcpy(nl.data(), h->d_nlist, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ns.data(), h->d_nshort, sizeof(int) * h->nclu, hipMemcpyDeviceToHost));
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
            const PairEntry &e = L[(size_t)c * h->list_cap + list_slot(k, ns[c], h->list_cap)];
            for (int lane = 0; lane < 64; ++lane) {
                if (!((e.mask >> lane) & 1ull)) continue;
                // the half list holds each pair once, in either orientation
                const int cj = (int)((unsigned)e.cj & ~kEntryWrap);
                const int u = ca[c * kClu + (lane >> 3)], v = ca[cj * kClu + (lane & 7)];
                const int ai = std::max(u, v), aj = std::min(u, v);
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

}  // extern "C"

// ---------------------------------------------------------------------
// Launch plans and batched force calls.

namespace {

template <typename T>
mdx_status kplan_upload(mdx_batch *b, KPlan<T> &k, bool maps) {
    if (k.h_args.empty()) return MDX_OK;
    if (k.h_args.size() > k.cap_args) {
        if (k.d_args) hipFree(k.d_args);
        k.d_args = nullptr;
        if (hipMalloc((void **)&k.d_args, sizeof(T) * k.h_args.size()) != hipSuccess) {
            b->err = "hipMalloc of plan arguments failed";
            return MDX_ERR_NOMEM;
        }
        k.cap_args = k.h_args.size();
    }
    if (hipMemcpy(k.d_args, k.h_args.data(), sizeof(T) * k.h_args.size(), hipMemcpyHostToDevice) != hipSuccess) {
        b->err = "plan upload failed";
        return MDX_ERR_HIP;
    }
    if (maps) {
        if (k.h_map.size() > k.cap_map) {
            if (k.d_map) hipFree(k.d_map);
            k.d_map = nullptr;
            if (hipMalloc((void **)&k.d_map, sizeof(int2) * k.h_map.size()) != hipSuccess) {
                b->err = "hipMalloc of plan block map failed";
                return MDX_ERR_NOMEM;
            }
            k.cap_map = k.h_map.size();
        }
        if (!k.h_map.empty() &&
            hipMemcpy(k.d_map, k.h_map.data(), sizeof(int2) * k.h_map.size(), hipMemcpyHostToDevice) != hipSuccess) {
            b->err = "plan upload failed";
            return MDX_ERR_HIP;
        }
    }
    return MDX_OK;
}

template <typename T>
void kplan_free(KPlan<T> &k) {
    if (k.d_args) hipFree(k.d_args);
    if (k.d_map) hipFree(k.d_map);
    k.d_args = nullptr;
    k.d_map = nullptr;
}

void drop_graphs(mdx_batch *b) {
    for (int g = 0; g < 3; ++g) {
        if (b->exec[g]) hipGraphExecDestroy(b->exec[g]);
        if (b->graph[g]) hipGraphDestroy(b->graph[g]);
        b->exec[g] = nullptr;
        b->graph[g] = nullptr;
    }
    if (b->md_exec) hipGraphExecDestroy(b->md_exec);
    if (b->md_graph) hipGraphDestroy(b->md_graph);
    b->md_exec = nullptr;
    b->md_graph = nullptr;
}

// Argument arrays of every kernel from the systems' current structs.  With
// maps = true the block maps are rebuilt too (first use, box changes).
mdx_status plan_fill(mdx_batch *b, bool maps) {
    const int K = (int)b->sys.size();
    b->pack.clear(); b->com.clear(); b->pairs.clear(); b->pairs64.clear(); b->special.clear();
    b->fur64.clear();
    b->rprep.clear(); b->sk.clear(); b->skred.clear(); b->rforce.clear();
    b->mz = 1;
    b->et.clear(); b->bonds.clear(); b->spc.clear(); b->angles.clear(); b->tors.clear(); b->terms.clear();
    b->springs.clear(); b->comb.clear(); b->red.clear(); b->redp.clear(); b->dred.clear();
    b->nmt_max = 1;
    b->nlj_max = 1;
    b->soft = false;
    b->pair_waves = b->sys.empty() ? 6 : b->sys[0]->pair_waves;
    for (int s = 0; s < K; ++s) {
        mdx_engine *h = b->sys[s];
        const int n = h->n;
        b->nmt_max = std::max(b->nmt_max, h->ntypes);
        if (h->d_dpart)
            b->dred.add(DetailArgs{h->det_rows, h->ndet, h->d_dpart, b->d_det + (size_t)s * kMaxDet}, h->ndet);
        b->nlj_max = std::max(b->nlj_max, h->nlj);
        b->soft = b->soft || h->p.ee_enabled;
        b->pack.add(h->a_pack, cdiv((long long)h->nclu_cap * kClu, 256));
        b->com.add(h->a_pack, h->ncomgrp);
        const bool fp64 = h->p.precision == MDX_PRECISION_FP64;
        if (fp64) {
            // j reactions by atomics into cleared words (cleared again by
            // the combine); FP64 FURIR
            b->pairs64.add(h->a_pair, h->nclu_cap);
            b->fur64.push_back(s);
        } else {
            b->pairs.add(h->a_pair, h->nclu_cap);
        }
        if (h->nspecial) b->special.add(h->a_special, cdiv(h->nspecial, 256));
        if (!fp64) {
            b->rprep.add(h->a_recip, cdiv(n, 256));
            b->sk.add(h->a_recip, h->nsgroups * h->nchunk);
            b->skred.add(h->a_recip, cdiv(h->npk + h->nfrows, kSkRedItems));
            b->rforce.add(h->a_recip, h->a_recip.natom_blocks * h->nslice);
            b->mz = std::max(b->mz, h->mz);
        }
        if (h->neterm) b->et.add(h->a_et, cdiv(h->neterm, 256));
        if (h->nb) b->bonds.add(h->a_bonds, cdiv(h->nb, 256));
        if (h->nspc) b->spc.add(h->a_spc, cdiv(h->nspc, 256));
        if (h->na) b->angles.add(h->a_angles, cdiv(h->na, 256));
        if (h->nt) b->tors.add(h->a_tors, cdiv(h->nt, 256));
        if (h->ntm) b->terms.add(h->a_terms, h->ntm);
        if (b->ring && K > 1) {
            mdx_engine *hp = b->sys[(s + 1) % K], *hm = b->sys[(s + K - 1) % K];
            SpringArgs sp;
            sp.n = n;
            sp.x = h->d_x; sp.y = h->d_y; sp.z = h->d_z;
            sp.xp = hp->d_x; sp.yp = hp->d_y; sp.zp = hp->d_z;
            sp.xm = hm->d_x; sp.ym = hm->d_y; sp.zm = hm->d_z;
            sp.mass = h->d_mass;
            sp.beadfac = b->beadfac;
            sp.nbeads = (double)K;
            sp.box = h->box;
            sp.fscale = h->fscale;
            sp.fx = h->d_fixsp; sp.fy = h->d_fixsp + n; sp.fz = h->d_fixsp + 2 * n;
            sp.epart = erow(h, h->row_spr);
            b->springs.add(sp, cdiv(n, 256));
        }
        CombineArgs ca;
        unsigned long long *fixG = h->d_fix, *fixHL = h->d_fix + 3 * (size_t)n, *fixHB = h->d_fix + 6 * (size_t)n;
        ca.n = n;
        ca.scale = (b->ring && K > 1) ? 1.0 / K : 1.0;
        ca.fsx = h->d_fs; ca.fsy = h->d_fs + n; ca.fsz = h->d_fs + 2 * n;
        ca.flx = h->d_fl; ca.fly = h->d_fl + n; ca.flz = h->d_fl + 2 * n;
        ca.recip = h->d_rpart;
        ca.nslice = fp64 ? (h->nkv > 0 ? 1 : 0) : h->nslice;
        ca.gx = fixG; ca.gy = fixG + n; ca.gz = fixG + 2 * n;
        ca.hlx = fixHL; ca.hly = fixHL + n; ca.hlz = fixHL + 2 * n;
        ca.hbx = fixHB; ca.hby = fixHB + n; ca.hbz = fixHB + 2 * n;
        ca.jf = h->d_fixj;
        ca.clear_jf = fp64 ? 1 : 0;
        ca.atom_slot = h->d_atom_slot;
        const bool spr = b->ring && K > 1;
        ca.spx = spr ? h->d_fixsp : nullptr;
        ca.spy = spr ? h->d_fixsp + n : nullptr;
        ca.spz = spr ? h->d_fixsp + 2 * n : nullptr;
        ca.inv_scale = 1.0 / h->fscale;
        ca.x = h->d_x; ca.y = h->d_y; ca.z = h->d_z;
        ca.atom_mol = h->d_mol;
        ca.cx = h->d_cx; ca.cy = h->d_cy; ca.cz = h->d_cz;
        ca.ox = h->d_out; ca.oy = h->d_out + n; ca.oz = h->d_out + 2 * n;
        ca.Gx = h->d_out + 3 * (size_t)n; ca.Gy = ca.Gx + n; ca.Gz = ca.Gx + 2 * n;
        ca.Hx = h->d_out + 6 * (size_t)n; ca.Hy = ca.Hx + n; ca.Hz = ca.Hx + 2 * n;
        ca.epart = erow(h, h->row_c);
        b->comb.add(ca, cdiv(n, 256));
        // energy rows: blocks of kRedRows rows, then one block per system
        // over their sums (one argument record per system, in order)
        const int nbp = cdiv(h->nclu_cap, kRedRows), nbr = cdiv(h->rows_main - h->row_pp, kRedRows);
        b->redp.add(ReduceArgs{h->nclu_cap, nbp, erow(h, h->row_pairs), erow(h, h->row_pp), erow(h, h->row_bp)},
                    nbp);
        b->red.add(ReduceArgs{h->rows_main - h->row_pp, nbr, erow(h, h->row_pp), b->d_energy + (size_t)s * Q_N,
                              erow(h, h->row_br)},
                   nbr);
    }
    CHK(kplan_upload(b, b->pack, maps)); CHK(kplan_upload(b, b->com, maps));
    CHK(kplan_upload(b, b->pairs, maps)); CHK(kplan_upload(b, b->pairs64, maps));
    CHK(kplan_upload(b, b->special, maps));
    CHK(kplan_upload(b, b->rprep, maps)); CHK(kplan_upload(b, b->sk, maps));
    CHK(kplan_upload(b, b->skred, maps)); CHK(kplan_upload(b, b->rforce, maps));
    CHK(kplan_upload(b, b->et, maps)); CHK(kplan_upload(b, b->bonds, maps)); CHK(kplan_upload(b, b->spc, maps));
    CHK(kplan_upload(b, b->angles, maps)); CHK(kplan_upload(b, b->tors, maps)); CHK(kplan_upload(b, b->terms, maps));
    CHK(kplan_upload(b, b->springs, maps)); CHK(kplan_upload(b, b->comb, maps));
    CHK(kplan_upload(b, b->red, maps)); CHK(kplan_upload(b, b->redp, maps));
    CHK(kplan_upload(b, b->dred, maps));
    b->seen.assign(K, 0);
    for (int s = 0; s < K; ++s) b->seen[s] = b->sys[s]->version;
    b->planned = true;
    return MDX_OK;
}

template <typename T, typename F>
void launch(mdx_batch *b, KPlan<T> &k, int nthreads, size_t lds, F kernel, hipStream_t st = nullptr) {
    if (k.nblocks() == 0) return;
    kernel<<<k.nblocks(), nthreads, lds, st ? st : b->stream>>>(k.d_args, k.d_map);
}

// Energy rows of every system: block sums, then one block per system.
void launch_reduce(mdx_batch *b, KPlan<ReduceArgs> &k, hipStream_t st) {
    if (k.nblocks() == 0) return;
    k_ereduce<<<k.nblocks(), 256, 0, st>>>(k.d_args, k.d_map);
    k_ereduce_last<<<(int)k.h_args.size(), 256, 0, st>>>(k.d_args);
}

void launch_forked(mdx_batch *b, int group, bool detail, bool terms);

// The kernel sequence of one force call over all systems of the batch.
// With stage timing the kernels run in order on the batch stream, with an
// event between stages; otherwise the call forks (launch_forked).
void launch_call(mdx_batch *b, int group, bool timing, bool detail, bool terms = false) {
    if (!timing) {
        launch_forked(b, group, detail, terms);
        return;
    }
    hipStream_t st = b->stream;
    auto mark = [&](int k) { if (timing) hipEventRecord(b->ev[k], st); };
    mark(0);
    launch(b, b->pack, 256, 0, k_pack);
    launch(b, b->com, 64, 0, k_com);
    mark(1);
    const size_t lj = sizeof(float2) * b->nlj_max * b->nlj_max;
    const size_t ldet = lj + sizeof(double) * (b->nmt_max + 1) * 2 * 64 + sizeof(int) * 8;
    if (b->pairs.nblocks()) {
        const int nb = b->pairs.nblocks();
        const PairArgs *pa = b->pairs.d_args;
        const int2 *pm = b->pairs.d_map;
        if (detail) {
            if (b->soft) k_pairs<true, true, 5><<<nb, 64, ldet, st>>>(pa, pm, group);
            else k_pairs<false, true, 5><<<nb, 64, ldet, st>>>(pa, pm, group);
        } else {
            if (b->soft) k_pairs<true, false, 5><<<nb, 64, lj, st>>>(pa, pm, group);
            else if (b->pair_waves == 5) k_pairs<false, false, 5><<<nb, 64, lj, st>>>(pa, pm, group);
            else k_pairs<false, false, 6><<<nb, 64, lj, st>>>(pa, pm, group);
        }
        k_jgather<<<nb, 64, 0, st>>>(pa, pm, group);
    }
    if (b->pairs64.nblocks()) {
        const int nb = b->pairs64.nblocks();
        const PairArgs *pa = b->pairs64.d_args;
        const int2 *pm = b->pairs64.d_map;
        const size_t l64 = sizeof(double) * (b->nmt_max + 1) * 2 * 64 + sizeof(int) * 8;
        if (detail) {
            if (b->soft) k_pairs64<true, true><<<nb, 64, l64, st>>>(pa, pm, group);
            else k_pairs64<false, true><<<nb, 64, l64, st>>>(pa, pm, group);
        } else {
            if (b->soft) k_pairs64<true, false><<<nb, 64, 0, st>>>(pa, pm, group);
            else k_pairs64<false, false><<<nb, 64, 0, st>>>(pa, pm, group);
        }
    }
    mark(2);
    if (b->special.nblocks()) {
        if (detail) k_special<true><<<b->special.nblocks(), 256, 0, st>>>(b->special.d_args, b->special.d_map, group);
        else k_special<false><<<b->special.nblocks(), 256, 0, st>>>(b->special.d_args, b->special.d_map, group);
    }
    mark(3);
    if (group != MDX_FORCES_FAST) {
        launch(b, b->rprep, 256, 0, k_recip_prep);
        mark(4);
        if (b->mz <= 16) launch(b, b->sk, 256, 0, k_sk<3>);
        else if (b->mz <= 32) launch(b, b->sk, 256, 0, k_sk<5>);
        else launch(b, b->sk, 256, 0, k_sk<7>);
        mark(5);
        launch(b, b->skred, 256, 0, k_sk_reduce);
        mark(6);
        if (b->mz <= 16) launch(b, b->rforce, 256, 0, k_recip_force<16>);
        else if (b->mz <= 32) launch(b, b->rforce, 256, 0, k_recip_force<32>);
        else launch(b, b->rforce, 256, 0, k_recip_force<48>);
        // FP64 precision mode: FURIR in FP64 into the force slice and the
        // first structure-factor energy row of the system
        for (int s : b->fur64) {
            mdx_engine *h = b->sys[s];
            const int n = h->n;
            if (h->nkv > 0)
                fur_launch(h->p, h->kbox, n, h->nkv, h->d_x, h->d_y, h->d_z, h->d_q, h->d_rpart, h->d_rpart + n,
                           h->d_rpart + 2 * (size_t)n, nullptr, erow(h, h->row_sk), h->d_furws, st);
        }
        mark(7);
        launch(b, b->et, 256, 0, k_eterms);
    } else {
        mark(4); mark(5); mark(6); mark(7);
    }
    if (group != MDX_FORCES_SLOW) {
        launch(b, b->spc, 256, 0, k_strbnd);
        launch(b, b->bonds, 256, 0, k_bonds);
        launch(b, b->angles, 256, 0, k_angles);
        launch(b, b->tors, 256, 0, k_tors);
        launch(b, b->springs, 256, 0, k_springs);
        if (terms) launch(b, b->terms, 256, 0, k_term_reduce);
    }
    mark(8);
    launch_reduce(b, b->redp, st);
    if (b->comb.nblocks()) k_combine<<<b->comb.nblocks(), 256, 0, st>>>(b->comb.d_args, b->comb.d_map, group);
    launch_reduce(b, b->red, st);
    hipMemcpyAsync(b->h_energy, b->d_energy, sizeof(double) * Q_N * b->sys.size(), hipMemcpyDeviceToHost, st);
    if (detail) {
        launch(b, b->dred, 256, 0, k_detail_reduce);
        hipMemcpyAsync(b->h_det, b->d_det, sizeof(double) * kMaxDet * b->sys.size(), hipMemcpyDeviceToHost, st);
    }
    mark(9);
}

// One force call as three chains that share nothing until the force
// combine: the batch stream runs the packing, the pair kernel, its gather
// and the sum of its energy rows; side stream 0 the reciprocal sum; side
// stream 1 the special pairs, exclusion corrections and bonded terms
// (integer atomics into accumulators that the previous call's combine left
// cleared).  The batch stream waits for both side streams before the
// combine.  Every sum has a fixed order or is an integer sum,
// so the results do not depend on how the chains interleave; the fork and
// join are events, so the sequence can be captured into a graph.
void launch_forked(mdx_batch *b, int group, bool detail, bool terms) {
    hipStream_t st = b->stream, sk = b->side[0], sb = b->side[1];
    hipEventRecord(b->fork_ev, st);
    hipStreamWaitEvent(sk, b->fork_ev, 0);
    hipStreamWaitEvent(sb, b->fork_ev, 0);
    // side stream 0: the reciprocal sum
    if (group != MDX_FORCES_FAST) {
        launch(b, b->rprep, 256, 0, k_recip_prep, sk);
        if (b->mz <= 16) launch(b, b->sk, 256, 0, k_sk<3>, sk);
        else if (b->mz <= 32) launch(b, b->sk, 256, 0, k_sk<5>, sk);
        else launch(b, b->sk, 256, 0, k_sk<7>, sk);
        launch(b, b->skred, 256, 0, k_sk_reduce, sk);
        if (b->mz <= 16) launch(b, b->rforce, 256, 0, k_recip_force<16>, sk);
        else if (b->mz <= 32) launch(b, b->rforce, 256, 0, k_recip_force<32>, sk);
        else launch(b, b->rforce, 256, 0, k_recip_force<48>, sk);
        for (int s : b->fur64) {
            mdx_engine *h = b->sys[s];
            const int n = h->n;
            if (h->nkv > 0)
                fur_launch(h->p, h->kbox, n, h->nkv, h->d_x, h->d_y, h->d_z, h->d_q, h->d_rpart, h->d_rpart + n,
                           h->d_rpart + 2 * (size_t)n, nullptr, erow(h, h->row_sk), h->d_furws, sk);
        }
    }
    // side stream 1: special pairs, exclusions, bonded terms, springs
    if (b->special.nblocks()) {
        if (detail) k_special<true><<<b->special.nblocks(), 256, 0, sb>>>(b->special.d_args, b->special.d_map, group);
        else k_special<false><<<b->special.nblocks(), 256, 0, sb>>>(b->special.d_args, b->special.d_map, group);
    }
    if (group != MDX_FORCES_FAST) launch(b, b->et, 256, 0, k_eterms, sb);
    if (group != MDX_FORCES_SLOW) {
        launch(b, b->spc, 256, 0, k_strbnd, sb);
        launch(b, b->bonds, 256, 0, k_bonds, sb);
        launch(b, b->angles, 256, 0, k_angles, sb);
        launch(b, b->tors, 256, 0, k_tors, sb);
        launch(b, b->springs, 256, 0, k_springs, sb);
        if (terms) launch(b, b->terms, 256, 0, k_term_reduce, sb);
    }
    // batch stream: packing, molecule centres, the pair stage
    launch(b, b->pack, 256, 0, k_pack);
    launch(b, b->com, 64, 0, k_com);
    const size_t lj = sizeof(float2) * b->nlj_max * b->nlj_max;
    const size_t ldet = lj + sizeof(double) * (b->nmt_max + 1) * 2 * 64 + sizeof(int) * 8;
    if (b->pairs.nblocks()) {
        const int nb = b->pairs.nblocks();
        const PairArgs *pa = b->pairs.d_args;
        const int2 *pm = b->pairs.d_map;
        if (detail) {
            if (b->soft) k_pairs<true, true, 5><<<nb, 64, ldet, st>>>(pa, pm, group);
            else k_pairs<false, true, 5><<<nb, 64, ldet, st>>>(pa, pm, group);
        } else {
            if (b->soft) k_pairs<true, false, 5><<<nb, 64, lj, st>>>(pa, pm, group);
            else if (b->pair_waves == 5) k_pairs<false, false, 5><<<nb, 64, lj, st>>>(pa, pm, group);
            else k_pairs<false, false, 6><<<nb, 64, lj, st>>>(pa, pm, group);
        }
        k_jgather<<<nb, 64, 0, st>>>(pa, pm, group);
    }
    if (b->pairs64.nblocks()) {
        const int nb = b->pairs64.nblocks();
        const PairArgs *pa = b->pairs64.d_args;
        const int2 *pm = b->pairs64.d_map;
        const size_t l64 = sizeof(double) * (b->nmt_max + 1) * 2 * 64 + sizeof(int) * 8;
        if (detail) {
            if (b->soft) k_pairs64<true, true><<<nb, 64, l64, st>>>(pa, pm, group);
            else k_pairs64<false, true><<<nb, 64, l64, st>>>(pa, pm, group);
        } else {
            if (b->soft) k_pairs64<true, false><<<nb, 64, 0, st>>>(pa, pm, group);
            else k_pairs64<false, false><<<nb, 64, 0, st>>>(pa, pm, group);
        }
    }
    launch_reduce(b, b->redp, st);
    // join
    hipEventRecord(b->join_ev[0], sk);
    hipEventRecord(b->join_ev[1], sb);
    hipStreamWaitEvent(st, b->join_ev[0], 0);
    hipStreamWaitEvent(st, b->join_ev[1], 0);
    if (b->comb.nblocks()) k_combine<<<b->comb.nblocks(), 256, 0, st>>>(b->comb.d_args, b->comb.d_map, group);
    launch_reduce(b, b->red, st);
    hipMemcpyAsync(b->h_energy, b->d_energy, sizeof(double) * Q_N * b->sys.size(), hipMemcpyDeviceToHost, st);
    if (detail) {
        launch(b, b->dred, 256, 0, k_detail_reduce);
        hipMemcpyAsync(b->h_det, b->d_det, sizeof(double) * kMaxDet * b->sys.size(), hipMemcpyDeviceToHost, st);
    }
}

// Checks and plan refresh before a force call (argument structs of systems
// whose lists or box changed are uploaded again; captured graphs are dropped
// when a launch shape changed).
mdx_status batch_prepare(mdx_batch *b, int group, bool detail) {
    if (group < 0 || group > 2) return MDX_ERR_ARG;
    if (detail)
        for (mdx_engine *h : b->sys) {
            mdx_status s = ensure_detail(h);
            if (s != MDX_OK) {
                b->err = h->err;
                return s;
            }
        }
    for (mdx_engine *h : b->sys)
        if (!h->have_list) {
            b->err = "mdx_compute before mdx_build_list";
            return MDX_ERR_STATE;
        }
    // work queued on the systems' own streams (coordinates) comes first
    for (mdx_engine *h : b->sys)
        if (h->stream != b->stream) hipStreamSynchronize(h->stream);
    bool stale = !b->planned;
    for (size_t s = 0; s < b->sys.size() && !stale; ++s) stale = b->seen[s] != b->sys[s]->version;
    if (stale) {
        // launch shapes depend on capacities only; a new box (k set) or a
        // regrown capacity changes the maps and the captured graphs
        std::vector<int> before = {b->pairs.nblocks(), b->pairs64.nblocks(), b->sk.nblocks(), b->rforce.nblocks(),
                                   b->skred.nblocks(),
                                   b->mz, b->nmt_max, b->nlj_max, b->pair_waves};
        CHK(plan_fill(b, true));
        std::vector<int> after = {b->pairs.nblocks(), b->pairs64.nblocks(), b->sk.nblocks(), b->rforce.nblocks(),
                                  b->skred.nblocks(),
                                  b->mz, b->nmt_max, b->nlj_max, b->pair_waves};
        if (before != after) drop_graphs(b);
    }
    return MDX_OK;
}

// Queue one force call for every system of the batch on the batch stream
// (no host synchronisation).  events: eager launches with stage events.
mdx_status batch_launch(mdx_batch *b, int group, bool detail, bool events, bool terms = false) {
    CHK(batch_prepare(b, group, detail));
    if (b->use_graph && !events && !detail && !terms) {
        if (!b->exec[group]) {
            HIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            launch_call(b, group, false, false);
            HIPCHK(b, hipStreamEndCapture(b->stream, &b->graph[group]));
            HIPCHK(b, hipGraphInstantiate(&b->exec[group], b->graph[group], nullptr, nullptr, 0));
        }
        HIPCHK(b, hipGraphLaunch(b->exec[group], b->stream));
    } else {
        launch_call(b, group, events, detail, terms);
    }
    return MDX_OK;
}

mdx_status batch_compute(mdx_batch *b, int group, bool detail, bool terms = false) {
    CHK(batch_launch(b, group, detail, b->timing, terms));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    HIPCHK(b, hipGetLastError());
    if (b->timing) {
        for (int k = ST_PREP; k <= ST_REDUCE; ++k) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, b->ev[k - 1], b->ev[k]);
            b->stage_ms[k] = ms;
        }
    }
    return MDX_OK;
}

void fill_energies(const mdx_batch *b, int s, int group, bool detail, mdx_energies *e) {
    const mdx_engine *h = b->sys[s];
    const double *q = b->h_energy + (size_t)s * Q_N;
    std::memset(e, 0, sizeof(*e));
    e->pe1 = q[Q_PE1]; e->pe2 = q[Q_PE2]; e->pels1 = q[Q_PELS1]; e->pels2 = q[Q_PELS2];
    e->spe = q[Q_SPE] + (group != MDX_FORCES_FAST ? h->spei : 0.0);
    e->pint = q[Q_PINT];
    e->vir1 = q[Q_VIR1]; e->vir2 = q[Q_VIR2]; e->virb = q[Q_VIRB];
    e->virx = q[Q_VIRX]; e->viry = q[Q_VIRY]; e->virz = q[Q_VIRZ];
    e->virfx = q[Q_VIRFX]; e->virfy = q[Q_VIRFY]; e->virfz = q[Q_VIRFZ];
    e->wirs = group != MDX_FORCES_FAST ? q[Q_WIRS] : 0.0;
    e->wirss = group != MDX_FORCES_SLOW ? q[Q_WIRSS] : 0.0;
    e->npairs_short = (int64_t)std::llround(q[Q_NPS]);
    e->npairs_long = (int64_t)std::llround(q[Q_NPL]);
    e->qek = q[Q_QEK];
    e->virpi = q[Q_VIRPI];
    for (int t = 0; t < h->ntypes && group != MDX_FORCES_FAST; ++t) e->spee[t] = h->spee_self[t];
    if (detail) {
        const double *d = b->h_det + (size_t)s * kMaxDet;
        const int ntp = h->ntypepairs, nmt = h->ntypes;
        for (int k = 0; k < ntp; ++k) { e->pote1[k] = d[k]; e->potl1[k] = d[ntp + k]; }
        for (int t = 0; t < nmt; ++t) { e->pes141[t] = d[2 * ntp + t]; e->psr141[t] = d[2 * ntp + nmt + t]; }
    }
}

mdx_status copy_forces(mdx_engine *h, double *fx, double *fy, double *fz, double *gx, double *gy, double *gz,
                       double *hx, double *hy, double *hz, hipMemcpyKind kind) {
    double *outs[9] = {fx, fy, fz, gx, gy, gz, hx, hy, hz};
    const size_t n = h->n, b = sizeof(double) * n;
    if (kind == hipMemcpyDeviceToHost) {
        // into pinned memory at full link speed, then to the caller's arrays
        CHK(ensure_stage(h));
        for (int k = 0; k < 9; ++k)
            if (outs[k]) HIPCHK(h, hipMemcpyAsync(h->h_stage + k * n, h->d_out + k * n, b, kind, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (int k = 0; k < 9; ++k)
            if (outs[k]) std::memcpy(outs[k], h->h_stage + k * n, b);
        return MDX_OK;
    }
    for (int k = 0; k < 9; ++k)
        if (outs[k]) HIPCHK(h, hipMemcpyAsync(outs[k], h->d_out + k * n, b, kind, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MDX_OK;
}


MdArgs md_args(mdx_batch *b, int s, double *trace, int cap) {
    mdx_engine *h = b->sys[s];
    const int n = h->n;
    MdArgs a{};
    a.n = n;
    a.nmol = h->nmol;
    a.x = h->d_x; a.y = h->d_y; a.z = h->d_z;
    a.vx = h->d_v; a.vy = h->d_v + n; a.vz = h->d_v + 2 * (size_t)n;
    a.ox = h->d_o; a.oy = h->d_o + n; a.oz = h->d_o + 2 * (size_t)n;
    a.ux = h->d_u; a.uy = h->d_u + n; a.uz = h->d_u + 2 * (size_t)n;
    a.fx = h->d_out; a.fy = h->d_out + n; a.fz = h->d_out + 2 * (size_t)n;
    a.gx = h->d_out + 3 * (size_t)n; a.gy = a.gx + n; a.gz = a.gx + 2 * (size_t)n;
    a.hx = h->d_out + 6 * (size_t)n; a.hy = a.hx + n; a.hz = a.hx + 2 * (size_t)n;
    a.minv = h->d_minv;
    // path-integral beads move with mass m / P (MASSD, setup.f:149)
    a.mfac = b->pimd ? 1.0 / (double)b->sys.size() : 1.0;
    a.imfac = b->pimd ? (double)b->sys.size() : 1.0;
    a.mass = h->d_mass;
    a.mstart = h->d_mstart;
    a.comgrp = h->d_comgrp;
    a.atom_mol = h->d_mol;
    a.cstart = h->d_cstart;
    a.cons = h->d_cons;
    a.cons_len2 = h->d_cons_len2;
    a.dt = h->md.dt;
    a.dts = h->md.integrator == MDX_MD_RESPA ? 0.5 * h->md.dt : h->md.dt;
    a.dti = h->md.dt / std::max(1, h->md.nfreq);
    a.accel = h->md.accel_factor;
    a.tol2 = h->md.shake_tol * h->md.shake_tol;
    a.bdtol = h->md.shake_bdtol;
    a.maxit = h->md.shake_maxit;
    a.nvt = h->md.tau_t > 0.0 ? 1 : 0;
    a.box = h->box;
    a.ekin_factor = h->md.ekin_factor;
    a.boltz = h->md.boltz;
    a.ndof = h->md.ndof;
    a.t0 = h->md.temperature;
    a.dqt = h->md.tau_t > 0.0 ? 1.0 / (h->md.tau_t * h->md.tau_t) : 0.0;
    a.state = h->d_mdstate;
    a.kepart = h->d_kepart;
    a.nke_blocks = cdiv(n, 256);
    a.energy = b->d_energy + (size_t)s * Q_N;
    a.spei = h->spei;
    a.trace = trace;
    a.trace_cap = cap;
    return a;
}

mdx_status md_plan(mdx_batch *b, int nsteps, bool want_trace) {
    const int K = (int)b->sys.size();
    if (want_trace && (size_t)K * nsteps * kTrace > b->trace_cap) {
        if (b->d_trace) hipFree(b->d_trace);
        b->d_trace = nullptr;
        b->trace_cap = 0;
        if (hipMalloc((void **)&b->d_trace, sizeof(double) * (size_t)K * nsteps * kTrace) != hipSuccess) {
            b->err = "hipMalloc of the MD trace failed";
            return MDX_ERR_NOMEM;
        }
        b->trace_cap = (size_t)K * nsteps * kTrace;
    }
    b->md_kd.clear(); b->md_sh.clear(); b->md_ke.clear(); b->md_th.clear(); b->md_gc.clear(); b->md_tr.clear();
    for (int s = 0; s < K; ++s) {
        mdx_engine *h = b->sys[s];
        const MdArgs a = md_args(b, s, want_trace ? b->d_trace + (size_t)s * nsteps * kTrace : nullptr, nsteps);
        b->md_kd.add(a, cdiv(h->n, 256));
        b->md_sh.add(a, cdiv(h->nmol, 64));
        b->md_ke.add(a, cdiv(h->n, 256));
        b->md_th.add(a, 1);
        b->md_gc.add(a, h->ncomgrp);
        b->md_tr.add(a, 1);
    }
    const void *before[6] = {b->md_kd.d_args, b->md_sh.d_args, b->md_ke.d_args, b->md_th.d_args, b->md_gc.d_args,
                             b->md_tr.d_args};
    CHK(kplan_upload(b, b->md_kd, true)); CHK(kplan_upload(b, b->md_sh, true));
    CHK(kplan_upload(b, b->md_ke, true)); CHK(kplan_upload(b, b->md_th, true));
    CHK(kplan_upload(b, b->md_gc, true)); CHK(kplan_upload(b, b->md_tr, true));
    const void *after[6] = {b->md_kd.d_args, b->md_sh.d_args, b->md_ke.d_args, b->md_th.d_args, b->md_gc.d_args,
                            b->md_tr.d_args};
    for (int k = 0; k < 6; ++k)
        if (before[k] != after[k] && b->md_exec) {
            hipGraphExecDestroy(b->md_exec);
            hipGraphDestroy(b->md_graph);
            b->md_exec = nullptr;
            b->md_graph = nullptr;
        }
    return MDX_OK;
}

// Bead-thermostat arguments of a path-integral ring batch.
PimdArgs pimd_args(mdx_batch *b) {
    const mdx_engine *h = b->sys[0];
    const double nb = (double)b->sys.size();
    PimdArgs a{};
    a.n = h->n;
    a.nbeads = (int)b->sys.size();
    a.nchain = std::max(1, b->pi.nchain);
    a.v = b->d_vptr;
    a.mass = h->d_mass;
    a.xi = b->d_xi;
    a.dti = h->md.dt / h->md.nfreq;
    const double tau = b->pi.tau_bead > 0.0 ? b->pi.tau_bead : 1.0;
    a.inv_tau2 = 1.0 / (tau * tau);
    a.fnh = a.inv_tau2 / (3.0 * nb);
    a.t0 = b->pi.temperature;
    a.conv = h->md.ekin_factor / (3.0 * h->md.boltz * nb * nb);
    a.nmol = h->nmol;
    a.mstart = h->d_mstart;
    a.x = b->d_vptr + a.nbeads;
    a.y = b->d_vptr + 2 * a.nbeads;
    a.z = b->d_vptr + 3 * a.nbeads;
    a.box = h->box;
    return a;
}

// nsteps long steps of the multiple-time-step algorithm DOUBLE
// (mdstep.f:281-530) for every system; a list rebuild falls before the
// force call of short step nfreq / 2 + 1 of every ICHNB-th long step
// (mdstep.f:389-390).  Without a rebuild the whole long step (nfreq force
// calls) is one captured hipGraph.
mdx_status batch_respa_run(mdx_batch *b, int nsteps, int nfreq, double *trace) {
    hipStream_t st = b->stream;
    const int nf2 = nfreq / 2 + 1;
    // a short step's kick and drift; path-integral beads take their
    // thermostat chains between the two (pimd.f:104-129)
    const PimdArgs pa = b->pimd ? pimd_args(b) : PimdArgs{};
    auto kick_drift = [&]() {
        if (!b->pimd) {
            launch(b, b->md_kd, 256, 0, k_md_kick_drift_fast);
            return;
        }
        launch(b, b->md_kd, 256, 0, k_md_kick_fast);
        if (b->pi.tau_bead > 0.0) k_pimd_beadnose<<<cdiv(pa.n, 256), 256, 0, st>>>(pa);
        launch(b, b->md_kd, 256, 0, k_md_drift);
    };
    // GETCOM; path-integral beads by their centroids' images
    auto getcom = [&]() {
        if (b->pimd) k_pimd_getcom<<<cdiv(pa.nmol, 256), 256, 0, st>>>(pa);
        else launch(b, b->md_gc, 64, 0, k_md_getcom);
    };
    auto scaling1 = [&]() {
        launch(b, b->md_ke, 256, 0, k_md_ke<false>);
        launch(b, b->md_th, 256, 0, k_md_thermo);
        launch(b, b->md_kd, 256, 0, k_md_scale);
    };
    auto scaling2 = [&]() {
        launch(b, b->md_ke, 256, 0, k_md_ke<true>);
        launch(b, b->md_th, 256, 0, k_md_thermo);
    };
    for (int step = 0; step < nsteps; ++step) {
        bool rebuild = false;
        for (mdx_engine *h : b->sys)
            rebuild = rebuild || (h->md.rebuild_every > 0 && (h->mstep + 1) % h->md.rebuild_every == 0);
        if (!rebuild && b->use_graph) {
            CHK(batch_prepare(b, MDX_FORCES_ALL, false));
            if (!b->md_exec) {
                HIPCHK(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                scaling1();
                launch(b, b->md_kd, 256, 0, k_md_kick_slow);
                for (int k = 1; k <= nfreq; ++k) {
                    kick_drift();
                    getcom();
                    launch_call(b, k < nfreq ? MDX_FORCES_FAST : MDX_FORCES_ALL, false, false);
                    launch(b, b->md_kd, 256, 0, k_md_kick_fast);
                }
                launch(b, b->md_kd, 256, 0, k_md_kick_slow);
                scaling2();
                launch(b, b->md_tr, 64, 0, k_md_trace);
                HIPCHK(b, hipStreamEndCapture(st, &b->md_graph));
                HIPCHK(b, hipGraphInstantiate(&b->md_exec, b->md_graph, nullptr, nullptr, 0));
            }
            HIPCHK(b, hipGraphLaunch(b->md_exec, st));
            for (mdx_engine *h : b->sys) h->mstep += 1;
            continue;
        }
        scaling1();
        launch(b, b->md_kd, 256, 0, k_md_kick_slow);
        for (int k = 1; k <= nfreq; ++k) {
            kick_drift();
            getcom();
            if (k == nf2) {
                bool synced = false;
                for (mdx_engine *h : b->sys)
                    if (h->md.rebuild_every > 0 && (h->mstep + 1) % h->md.rebuild_every == 0) {
                        if (!synced) {
                            HIPCHK(b, hipStreamSynchronize(st));
                            synced = true;
                        }
                        mdx_status s = mdx_build_list(h);
                        if (s != MDX_OK) {
                            b->err = h->err;
                            return s;
                        }
                    }
            }
            CHK(batch_launch(b, k < nfreq ? MDX_FORCES_FAST : MDX_FORCES_ALL, false, false));
            launch(b, b->md_kd, 256, 0, k_md_kick_fast);
        }
        launch(b, b->md_kd, 256, 0, k_md_kick_slow);
        scaling2();
        launch(b, b->md_tr, 64, 0, k_md_trace);
        for (mdx_engine *h : b->sys) h->mstep += 1;
    }
    HIPCHK(b, hipStreamSynchronize(st));
    HIPCHK(b, hipGetLastError());
    if (trace && nsteps > 0)
        HIPCHK(b, hipMemcpy(trace, b->d_trace, sizeof(double) * b->sys.size() * nsteps * kTrace, hipMemcpyDeviceToHost));
    return MDX_OK;
}

// nsteps leap-frog steps of every system; lists are rebuilt (on the
// systems' own streams, after the batch stream drains) when a system's
// MSTEP reaches a multiple of its ICHNB.
mdx_status batch_md_run(mdx_batch *b, int nsteps, double *trace) {
    if (nsteps < 0) return MDX_ERR_ARG;
    for (mdx_engine *h : b->sys)
        if (!h->md_on) {
            b->err = "mdx_md_run before mdx_md_init";
            return MDX_ERR_STATE;
        }
    for (mdx_engine *h : b->sys)
        if (h->stream != b->stream) hipStreamSynchronize(h->stream);
    const int integrator = b->sys[0]->md.integrator, nfreq = b->sys[0]->md.nfreq;
    for (mdx_engine *h : b->sys)
        if (h->md.integrator != integrator || (integrator == MDX_MD_RESPA && h->md.nfreq != nfreq)) {
            b->err = "the systems of a batch must share the integrator and nfreq";
            return MDX_ERR_ARG;
        }
    CHK(md_plan(b, std::max(nsteps, 1), trace != nullptr));
    const int kind = integrator * 65536 + (integrator != MDX_MD_LEAPFROG ? nfreq : 0);
    if (b->md_exec && b->md_exec_kind != kind) {
        hipGraphExecDestroy(b->md_exec);
        hipGraphDestroy(b->md_graph);
        b->md_exec = nullptr;
        b->md_graph = nullptr;
    }
    b->md_exec_kind = kind;
    for (mdx_engine *h : b->sys)
        HIPCHK(b, hipMemsetAsync(h->d_mdstate + MD_STEP, 0, sizeof(double), b->stream));
    if (integrator == MDX_MD_RESPA) return batch_respa_run(b, nsteps, nfreq, trace);
    if (integrator == MDX_MD_PIMD) {
        if (!b->pimd) {
            b->err = "path-integral MD before mdx_batch_pimd_init";
            return MDX_ERR_STATE;
        }
        return batch_respa_run(b, nsteps, nfreq, trace);
    }
    hipStream_t st = b->stream;
    auto md_kernels = [&]() {
        launch(b, b->md_kd, 256, 0, k_md_kick_drift);
        launch(b, b->md_sh, 64, 0, k_md_shake);
        launch(b, b->md_ke, 256, 0, k_md_ke<true>);
        launch(b, b->md_th, 256, 0, k_md_thermo);
        launch(b, b->md_gc, 64, 0, k_md_getcom);
    };
    for (int step = 0; step < nsteps; ++step) {
        bool rebuild = false;
        for (mdx_engine *h : b->sys)
            rebuild = rebuild || (h->md.rebuild_every > 0 && (h->mstep + 1) % h->md.rebuild_every == 0);
        if (!rebuild && b->use_graph) {
            // the whole step (integration, force call, trace) as one graph
            CHK(batch_prepare(b, MDX_FORCES_ALL, false));
            if (!b->md_exec) {
                HIPCHK(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                md_kernels();
                launch_call(b, MDX_FORCES_ALL, false, false);
                launch(b, b->md_tr, 64, 0, k_md_trace);
                HIPCHK(b, hipStreamEndCapture(st, &b->md_graph));
                HIPCHK(b, hipGraphInstantiate(&b->md_exec, b->md_graph, nullptr, nullptr, 0));
            }
            HIPCHK(b, hipGraphLaunch(b->md_exec, st));
            for (mdx_engine *h : b->sys) h->mstep += 1;
            continue;
        }
        md_kernels();
        bool synced = false;
        for (mdx_engine *h : b->sys) {
            h->mstep += 1;
            if (h->md.rebuild_every > 0 && h->mstep % h->md.rebuild_every == 0) {
                if (!synced) {
                    HIPCHK(b, hipStreamSynchronize(st));
                    synced = true;
                }
                mdx_status s = mdx_build_list(h);
                if (s != MDX_OK) {
                    b->err = h->err;
                    return s;
                }
            }
        }
        CHK(batch_launch(b, MDX_FORCES_ALL, false, false));
        launch(b, b->md_tr, 64, 0, k_md_trace);
    }
    HIPCHK(b, hipStreamSynchronize(st));
    HIPCHK(b, hipGetLastError());
    if (trace && nsteps > 0)
        HIPCHK(b, hipMemcpy(trace, b->d_trace, sizeof(double) * b->sys.size() * nsteps * kTrace, hipMemcpyDeviceToHost));
    return MDX_OK;
}

}  // namespace

extern "C" {

mdx_status mdx_compute(mdx_engine *h, int32_t group, int32_t want, double *fx, double *fy, double *fz, double *gx,
                       double *gy, double *gz, double *hx, double *hy, double *hz, mdx_energies *e) {
    if (!h) return MDX_ERR_ARG;
    const bool detail = (want & MDX_WANT_DETAIL) != 0, terms = (want & MDX_WANT_TERMS) != 0;
    mdx_batch *b = h->self;
    b->timing = h->timing;
    mdx_status s = batch_compute(b, group, detail, terms);
    if (s != MDX_OK) {
        h->err = b->err;
        return s;
    }
    for (int k = ST_PREP; k < ST_N; ++k) h->stage_ms[k] = b->stage_ms[k];
    if (e) fill_energies(b, 0, group, detail, e);
    return copy_forces(h, fx, fy, fz, gx, gy, gz, hx, hy, hz, hipMemcpyDeviceToHost);
}

mdx_status mdx_compute_device(mdx_engine *h, int32_t group, int32_t want, double *dfx, double *dfy, double *dfz,
                              mdx_energies *e) {
    if (!h) return MDX_ERR_ARG;
    const bool detail = (want & MDX_WANT_DETAIL) != 0, terms = (want & MDX_WANT_TERMS) != 0;
    mdx_batch *b = h->self;
    b->timing = h->timing;
    mdx_status s = batch_compute(b, group, detail, terms);
    if (s != MDX_OK) {
        h->err = b->err;
        return s;
    }
    for (int k = ST_PREP; k < ST_N; ++k) h->stage_ms[k] = b->stage_ms[k];
    if (e) fill_energies(b, 0, group, detail, e);
    return copy_forces(h, dfx, dfy, dfz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       hipMemcpyDeviceToDevice);
}

mdx_status mdx_get_forces(mdx_engine *h, double *fx, double *fy, double *fz, double *gx, double *gy, double *gz,
                          double *hx, double *hy, double *hz) {
    if (!h) return MDX_ERR_ARG;
    return copy_forces(h, fx, fy, fz, gx, gy, gz, hx, hy, hz, hipMemcpyDeviceToHost);
}

// The cluster grid with column width scale `scale`: cluster arrays, list
// storage and energy rows allocated again (capacities follow the grid);
// the list must be built again before the next force call.
static mdx_status regrid_clusters(mdx_engine *h, double scale) {
    h->cw_scale = scale;
    CHK(alloc_clusters(h));
    layout_rows(h);
    CHK(dalloc_zero(h, &h->d_epart, (size_t)h->rows_total * Q_N));
    h->have_list = false;
    fill_args(h);
    return MDX_OK;
}

mdx_status mdx_set_cluster_grid(mdx_engine *h, double scale) {
    if (!h || !(scale >= 0.5 && scale <= 2.0)) return MDX_ERR_ARG;
    const bool had = h->have_list;
    CHK(regrid_clusters(h, scale));
    if (had) CHK(mdx_build_list(h));
    return MDX_OK;
}

mdx_status mdx_set_pair_waves(mdx_engine *h, int32_t waves) {
    if (!h || (waves != 5 && waves != 6)) return MDX_ERR_ARG;
    h->pair_waves = waves;
    h->version++;                    // batches plan and capture again
    return MDX_OK;
}

mdx_status mdx_pair_waves(const mdx_engine *h, int32_t *waves) {
    if (!h || !waves) return MDX_ERR_ARG;
    *waves = h->pair_waves;
    return MDX_OK;
}

mdx_status mdx_cluster_grid(const mdx_engine *h, double *column_width, int32_t *ncx, int32_t *ncy) {
    if (!h) return MDX_ERR_ARG;
    if (column_width) *column_width = h->p.box[0] / std::max(1, h->ncx);
    if (ncx) *ncx = h->ncx;
    if (ncy) *ncy = h->ncy;
    return MDX_OK;
}

mdx_status mdx_autotune(mdx_engine *h, int32_t *chunk_out, int32_t *nslice_out) {
    if (!h) return MDX_ERR_ARG;
    if (!h->have_list) {
        h->err = "mdx_autotune before mdx_build_list";
        return MDX_ERR_STATE;
    }
    if (h->p.precision == MDX_PRECISION_FP64) {
        // the FP64 reciprocal sum has no tunable shape
        if (chunk_out) *chunk_out = h->chunk;
        if (nslice_out) *nslice_out = h->nslice;
        return MDX_OK;
    }
    mdx_batch *b = h->self;
    const bool timing0 = b->timing;
    b->timing = true;
    auto trial = [&](int chunk, int nslice, double &sk_ms, double &force_ms) -> mdx_status {
        CHK(set_kspace_shape(h, chunk, nslice));
        sk_ms = force_ms = 1e30;
        for (int rep = 0; rep < 4; ++rep) {
            mdx_status st = batch_compute(b, MDX_FORCES_SLOW, false);
            if (st != MDX_OK) {
                h->err = b->err;
                return st;
            }
            if (rep == 0) continue;     // warm-up (plan, allocation)
            sk_ms = std::min(sk_ms, b->stage_ms[ST_SK] + b->stage_ms[ST_SKRED]);
            force_ms = std::min(force_ms, b->stage_ms[ST_RFORCE] + b->stage_ms[ST_REDUCE]);
        }
        return MDX_OK;
    };
    const int wav = cdiv(h->n, 16 * kForceTiles);
    int best_chunk = h->chunk, best_slice = h->nslice;
    double best_sk = 1e30, best_force = 1e30, sk = 0.0, fo = 0.0;
    for (int chunk : {512, 1024, 2048, 4096}) {
        if (chunk > 2 * h->n && chunk != 512) continue;
        CHK(trial(chunk, best_slice, sk, fo));
        if (sk < best_sk) { best_sk = sk; best_chunk = chunk; }
    }
    for (int waves : {2048, 3072, 4096, 6144, 8192, 12288, 16384}) {
        const int ns = std::max(1, std::min(64, cdiv(waves, std::max(1, wav))));
        if (ns == best_slice && best_force < 1e30) continue;
        CHK(trial(best_chunk, ns, sk, fo));
        if (fo < best_force) { best_force = fo; best_slice = ns; }
    }
    CHK(set_kspace_shape(h, best_chunk, best_slice));
    // Cluster geometry: the width of the x-y columns whose z-sorted atoms
    // form the 8-atom clusters.  Wider columns give flatter clusters,
    // narrower ones taller clusters; the column count also sets which
    // clusters share a column and how the i-clusters' blocks fall on the
    // XCDs, and the pair stage's time is not smooth in it (at 96,174
    // atoms: 22 columns along x 0.27 ms, 21 or 24 columns 0.24 ms).  Every
    // column count along x from 0.75 to 1.3 times the cube root of 8
    // atoms' volume is tried, scored by the pair stage plus a tenth of a
    // list build (a rebuild every ten steps); the list is rebuilt at the
    // current coordinates.
    const bool htiming0 = h->timing;
    h->timing = true;
    double best_cw = h->cw_scale, best_score = 1e30;
    const double cw0 = std::cbrt(kClu * h->p.box[0] * h->p.box[1] * h->p.box[2] / h->n);
    std::vector<double> scales;
    for (int nc = std::max(1, (int)std::floor(h->p.box[0] / (1.3 * cw0)));
         nc <= (int)std::ceil(h->p.box[0] / (0.75 * cw0)); ++nc) {
        // exactly nc columns along x (the grid takes floor(L / width))
        const double sc = h->p.box[0] / (nc * cw0) * (1.0 - 1e-9);
        if (sc >= 0.5 && sc <= 2.0) scales.push_back(sc);
    }
    for (double cw : scales) {
        CHK(regrid_clusters(h, cw));
        CHK(mdx_build_list(h));
        const double list_ms = h->stage_ms[ST_LIST];
        double pair_ms = 1e30;
        for (int rep = 0; rep < 6; ++rep) {
            mdx_status st = batch_compute(b, MDX_FORCES_ALL, false);
            if (st != MDX_OK) {
                h->err = b->err;
                return st;
            }
            if (rep > 0) pair_ms = std::min(pair_ms, (double)b->stage_ms[ST_PAIRS]);
        }
        const double score = pair_ms + 0.1 * list_ms;
        if (score < best_score) { best_score = score; best_cw = cw; }
    }
    CHK(regrid_clusters(h, best_cw));
    CHK(mdx_build_list(h));
    // the plain pair kernel's wavefronts per SIMD on the chosen grid
    int best_pw = h->pair_waves;
    double best_pair = 1e30;
    for (int pw : {6, 5}) {
        CHK(mdx_set_pair_waves(h, pw));
        double pair_ms = 1e30;
        for (int rep = 0; rep < 6; ++rep) {
            mdx_status st = batch_compute(b, MDX_FORCES_ALL, false);
            if (st != MDX_OK) {
                h->err = b->err;
                return st;
            }
            if (rep > 0) pair_ms = std::min(pair_ms, (double)b->stage_ms[ST_PAIRS]);
        }
        if (pair_ms < best_pair) { best_pair = pair_ms; best_pw = pw; }
    }
    CHK(mdx_set_pair_waves(h, best_pw));
    h->timing = htiming0;
    b->timing = timing0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[tune_key(h)] = TuneShape{best_chunk, best_slice, best_cw, best_pw};
    }
    if (chunk_out) *chunk_out = best_chunk;
    if (nslice_out) *nslice_out = best_slice;
    return MDX_OK;
}

static mdx_status export_kvectors(const std::vector<int4> &ks, int32_t *nx, int32_t *ny, int32_t *nz, int64_t cap,
                                  int64_t *nkv) {
    if (!nkv || cap < 0 || (cap > 0 && (!nx || !ny || !nz))) return MDX_ERR_ARG;
    *nkv = (int64_t)ks.size();
    if ((int64_t)ks.size() > cap && cap > 0) return MDX_ERR_ARG;
    for (size_t k = 0; k < ks.size() && (int64_t)k < cap; ++k) {
        nx[k] = ks[k].x;
        ny[k] = ks[k].y;
        nz[k] = ks[k].z;
    }
    return MDX_OK;
}

mdx_status mdx_get_kvectors(const mdx_engine *h, int32_t *nx, int32_t *ny, int32_t *nz, int64_t cap,
                            int64_t *nkv) {
    if (!h) return MDX_ERR_ARG;
    return export_kvectors(h->h_kidx, nx, ny, nz, cap, nkv);
}

mdx_status mdx_enumerate_kvectors(const mdx_params *p, int32_t *nx, int32_t *ny, int32_t *nz, int64_t cap,
                                  int64_t *nkv) {
    if (!p || !(p->alpha > 0.0) || !(p->fexp > 0.0)) return MDX_ERR_ARG;
    for (int k = 0; k < 3; ++k)
        if (!(p->box[k] > 0.0)) return MDX_ERR_ARG;
    std::vector<int4> ks;
    enumerate_k(*p, ks);
    return export_kvectors(ks, nx, ny, nz, cap, nkv);
}

mdx_status mdx_furir_workspace_bytes(const mdx_params *p, int32_t natoms, int64_t *bytes) {
    if (!fur_params_ok(p) || natoms < 0 || !bytes) return MDX_ERR_ARG;
    const long long nkv = count_k(*p);
    if (nkv > 4000000) return MDX_ERR_NKV;
    *bytes = (int64_t)fur_layout(natoms, (int)nkv).total;
    return MDX_OK;
}

mdx_status mdx_furir_device(const mdx_params *p, int32_t natoms, const double *x, const double *y, const double *z,
                            const double *q, double *fx, double *fy, double *fz, double *energy4, void *workspace,
                            int64_t workspace_bytes, void *stream) {
    if (!fur_params_ok(p) || natoms < 0 || !energy4 || !workspace) return MDX_ERR_ARG;
    if (natoms > 0 && (!x || !y || !z || !q || !fx || !fy || !fz)) return MDX_ERR_ARG;
    const long long nkv = count_k(*p);
    if (nkv > 4000000) return MDX_ERR_NKV;
    if ((int64_t)fur_layout(natoms, (int)nkv).total > workspace_bytes) return MDX_ERR_ARG;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    fur_launch(*p, nullptr, natoms, nkv, x, y, z, q, fx, fy, fz, energy4, nullptr, static_cast<char *>(workspace), st);
    return hipGetLastError() == hipSuccess ? MDX_OK : MDX_ERR_HIP;
}

}  // extern "C"

// Scratch of mdx_furir_device_noalloc: one array per GPU in the code
// object's device memory.  In host code `scratch` names a host-side stand-in,
// not a device address: kernels get the address hipGetSymbolAddress returns
// for the current device, looked up once per device on its first call.
static __device__ __align__(256) unsigned char scratch[MDX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

mdx_status scratch_of_current_device(char **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return MDX_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MDX_ERR_HIP;
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<char *>(g_scratch_ptr[dev]);
    return MDX_OK;
}
}  // namespace

extern "C" {

mdx_status mdx_furir_device_noalloc(const mdx_params *p, int32_t natoms, const double *x, const double *y,
                                    const double *z, const double *q, double *fx, double *fy, double *fz,
                                    double *energy4, void *stream) {
    if (!fur_params_ok(p) || natoms < 0 || !energy4) return MDX_ERR_ARG;
    if (natoms > 0 && (!x || !y || !z || !q || !fx || !fy || !fz)) return MDX_ERR_ARG;
    const long long nkv = count_k(*p);
    if (nkv > 4000000) return MDX_ERR_NKV;
    if ((long long)fur_layout(natoms, (int)nkv).total > MDX_NOALLOC_SCRATCH_BYTES) return MDX_ERR_LIMIT;
    char *ws = nullptr;
    CHK(scratch_of_current_device(&ws));
    fur_launch(*p, nullptr, natoms, nkv, x, y, z, q, fx, fy, fz, energy4, nullptr, ws, reinterpret_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? MDX_OK : MDX_ERR_HIP;
}

mdx_status mdx_set_kspace_shape(mdx_engine *h, int32_t chunk, int32_t nslice) {
    if (!h || chunk <= 0 || nslice <= 0) return MDX_ERR_ARG;
    return set_kspace_shape(h, chunk, nslice);
}

mdx_status mdx_set_timing(mdx_engine *h, int32_t on) {
    if (!h) return MDX_ERR_ARG;
    h->timing = on != 0;
    h->self->use_graph = !h->timing;
    return MDX_OK;
}

mdx_status mdx_stage_times(mdx_engine *h, double *ms, int32_t n) {
    if (!h || !ms) return MDX_ERR_ARG;
    for (int k = 0; k < n && k < ST_N; ++k) ms[k] = h->stage_ms[k];
    return MDX_OK;
}

mdx_status mdx_memory_bytes(const mdx_engine *h, int64_t *bytes) {
    if (!h || !bytes) return MDX_ERR_ARG;
    int64_t t = 0;
    for (auto &kv : h->dev_sizes) t += (int64_t)kv.second;
    *bytes = t;
    return MDX_OK;
}

mdx_status mdx_md_init(mdx_engine *h, const mdx_md_params *p, const double *vx, const double *vy,
                       const double *vz) {
    if (!h || !p || !vx || !vy || !vz) return MDX_ERR_ARG;
    if (!h->have_pos) {
        h->err = "mdx_md_init before mdx_set_positions";
        return MDX_ERR_STATE;
    }
    if (!(p->dt > 0.0) || !(p->ndof > 0.0) || !(p->boltz > 0.0) || p->ncons < 0 ||
        (p->ncons > 0 && (!p->cons_i || !p->cons_j || !p->cons_len)) || p->shake_maxit < 0 ||
        (p->integrator != MDX_MD_LEAPFROG && p->integrator != MDX_MD_RESPA && p->integrator != MDX_MD_PIMD) ||
        (p->integrator != MDX_MD_LEAPFROG && (p->nfreq < 1 || p->nfreq > 1000 || p->ncons > 0)) ||
        (p->integrator == MDX_MD_PIMD && p->tau_t > 0.0)) {
        h->err = "invalid MD parameters";
        return MDX_ERR_ARG;
    }
    const int n = h->n;
    // constraints grouped by molecule, both atoms in it, at most 64 atoms
    std::vector<int> cstart(h->nmol + 1, 0);
    std::vector<int2> cons(std::max(1, p->ncons));
    std::vector<double> len2(std::max(1, p->ncons));
    int prev = -1;
    for (int c = 0; c < p->ncons; ++c) {
        const int i = p->cons_i[c], j = p->cons_j[c];
        if (i < 0 || i >= n || j < 0 || j >= n || i == j) return h->err = "constraint atom out of range", MDX_ERR_ARG;
        const int m = h->h_mol[i];
        if (h->h_mol[j] != m || m < prev) return h->err = "constraints must be grouped by molecule", MDX_ERR_ARG;
        if (h->h_mstart[m + 1] - h->h_mstart[m] > 64)
            return h->err = "constrained molecules may have at most 64 atoms", MDX_ERR_ARG;
        prev = m;
        cstart[m + 1] += 1;
        cons[c] = make_int2(i, j);
        len2[c] = p->cons_len[c] * p->cons_len[c];
    }
    for (int m = 0; m < h->nmol; ++m) cstart[m + 1] += cstart[m];
    std::vector<double> minv(n);
    for (int i = 0; i < n; ++i) minv[i] = 1.0 / h->h_mass[i];
    std::vector<double> v(3 * (size_t)n);
    std::memcpy(v.data(), vx, sizeof(double) * n);
    std::memcpy(v.data() + n, vy, sizeof(double) * n);
    std::memcpy(v.data() + 2 * (size_t)n, vz, sizeof(double) * n);
    std::vector<double> state(MD_NSTATE, 0.0);
    state[MD_SC] = p->sc0;
    mdx_status s;
    if ((s = upload(h, &h->d_v, v.data(), v.size())) || (s = dalloc(h, &h->d_o, 3 * (size_t)n)) ||
        (s = dalloc(h, &h->d_u, 3 * (size_t)n)) || (s = upload(h, &h->d_minv, minv.data(), n)) ||
        (s = upload(h, &h->d_mdstate, state.data(), MD_NSTATE)) || (s = dalloc(h, &h->d_kepart, cdiv(n, 256))) ||
        (s = upload(h, &h->d_cstart, cstart.data(), cstart.size())) ||
        (s = upload(h, &h->d_cons, cons.data(), cons.size())) ||
        (s = upload(h, &h->d_cons_len2, len2.data(), len2.size())))
        return s;
    h->md = *p;
    h->md.cons_i = h->md.cons_j = nullptr;
    h->md.cons_len = nullptr;
    h->ncons = p->ncons;
    h->mstep = p->step0;
    h->md_on = true;
    // SULTAN's and DOUBLE's start: list and forces at the initial positions
    CHK(mdx_build_list(h));
    mdx_batch *b = h->self;
    b->timing = false;
    s = batch_compute(b, MDX_FORCES_ALL, false);
    b->timing = h->timing;
    if (s != MDX_OK) h->err = b->err;
    return s;
}

mdx_status mdx_md_run(mdx_engine *h, int32_t nsteps, double *trace) {
    if (!h) return MDX_ERR_ARG;
    mdx_status s = batch_md_run(h->self, nsteps, trace);
    if (s != MDX_OK) h->err = h->self->err;
    return s;
}

mdx_status mdx_batch_md_run(mdx_batch *b, int32_t nsteps, double *trace) {
    if (!b) return MDX_ERR_ARG;
    return batch_md_run(b, nsteps, trace);
}

mdx_status mdx_batch_pimd_init(mdx_batch *b, const mdx_pimd_params *p) {
    if (!b || !p) return MDX_ERR_ARG;
    const int K = (int)b->sys.size();
    if (!b->ring || K < 2) {
        b->err = "path-integral MD needs a ring batch of at least 2 beads";
        return MDX_ERR_STATE;
    }
    const mdx_engine *h0 = b->sys[0];
    for (mdx_engine *h : b->sys)
        if (!h->md_on || h->md.integrator != MDX_MD_PIMD || h->md.nfreq != h0->md.nfreq || h->n != h0->n) {
            b->err = "every bead needs mdx_md_init with MDX_MD_PIMD and the same nfreq";
            return MDX_ERR_STATE;
        }
    if (p->tau_bead < 0.0 || (p->tau_bead > 0.0 && (p->nchain < 1 || !(p->temperature > 0.0)))) {
        b->err = "invalid bead thermostat";
        return MDX_ERR_ARG;
    }
    const int n = h0->n, nchain = std::max(1, p->nchain);
    if (b->d_xi) hipFree(b->d_xi);
    if (b->d_vptr) hipFree(b->d_vptr);
    b->d_xi = nullptr;
    b->d_vptr = nullptr;
    std::vector<double *> vp(4 * (size_t)K);
    for (int s = 0; s < K; ++s) {
        vp[s] = b->sys[s]->d_v;
        vp[K + s] = b->sys[s]->d_x;
        vp[2 * K + s] = b->sys[s]->d_y;
        vp[3 * K + s] = b->sys[s]->d_z;
    }
    if (hipMalloc((void **)&b->d_xi, sizeof(double) * (size_t)nchain * n) != hipSuccess ||
        hipMalloc((void **)&b->d_vptr, sizeof(double *) * 4 * K) != hipSuccess) {
        b->err = "device allocation failed";
        return MDX_ERR_NOMEM;
    }
    HIPCHK(b, hipMemsetAsync(b->d_xi, 0, sizeof(double) * (size_t)nchain * n, b->stream));
    HIPCHK(b, hipMemcpyAsync(b->d_vptr, vp.data(), sizeof(double *) * vp.size(), hipMemcpyHostToDevice, b->stream));
    b->pi = *p;
    b->pi.nchain = nchain;
    b->pimd = true;
    if (b->md_exec) {
        hipGraphExecDestroy(b->md_exec);
        hipGraphDestroy(b->md_graph);
        b->md_exec = nullptr;
        b->md_graph = nullptr;
    }
    // the ring's forces at the start positions (pimd.f:46-69)
    for (mdx_engine *h : b->sys)
        if (h->stream != b->stream) HIPCHK(b, hipStreamSynchronize(h->stream));
    CHK(batch_compute(b, MDX_FORCES_ALL, false));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return MDX_OK;
}

mdx_status mdx_batch_pimd_chains(mdx_batch *b, double *xi) {
    if (!b || !xi) return MDX_ERR_ARG;
    if (!b->pimd) {
        b->err = "mdx_batch_pimd_chains before mdx_batch_pimd_init";
        return MDX_ERR_STATE;
    }
    const size_t cnt = (size_t)b->pi.nchain * b->sys[0]->n;
    HIPCHK(b, hipMemcpyAsync(xi, b->d_xi, sizeof(double) * cnt, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return MDX_OK;
}

mdx_status mdx_md_get_state(mdx_engine *h, double *x, double *y, double *z, double *vx, double *vy, double *vz) {
    if (!h) return MDX_ERR_ARG;
    if (!h->md_on) {
        h->err = "mdx_md_get_state before mdx_md_init";
        return MDX_ERR_STATE;
    }
    const size_t b = sizeof(double) * h->n;
    const double *src[6] = {h->d_x, h->d_y, h->d_z, h->d_v, h->d_v + h->n, h->d_v + 2 * (size_t)h->n};
    double *dst[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) HIPCHK(h, hipMemcpyAsync(dst[k], src[k], b, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MDX_OK;
}

mdx_status mdx_md_shake_failures(mdx_engine *h, int64_t *count) {
    if (!h || !count) return MDX_ERR_ARG;
    if (!h->md_on) {
        h->err = "mdx_md_shake_failures before mdx_md_init";
        return MDX_ERR_STATE;
    }
    double v = 0.0;
    HIPCHK(h, hipMemcpyAsync(&v, h->d_mdstate + MD_SHAKE_FAIL, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *count = (int64_t)v;
    return MDX_OK;
}

mdx_status mdx_batch_create(mdx_engine *const *systems, int32_t k, mdx_batch **out) {
    if (!systems || k <= 0 || !out) return MDX_ERR_ARG;
    mdx_batch *b = new mdx_batch();
    for (int s = 0; s < k; ++s) {
        if (!systems[s]) {
            delete b;
            return MDX_ERR_ARG;
        }
        b->sys.push_back(systems[s]);
    }
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return MDX_ERR_HIP;
    }
    b->own_stream = true;
    for (int e = 0; e < 12; ++e) hipEventCreate(&b->ev[e]);
    for (int k = 0; k < 2; ++k)
        if (hipStreamCreate(&b->side[k]) != hipSuccess ||
            hipEventCreateWithFlags(&b->join_ev[k], hipEventDisableTiming) != hipSuccess) {
            *out = b;
            return MDX_ERR_HIP;
        }
    if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
        *out = b;
        return MDX_ERR_HIP;
    }
    if (hipMalloc((void **)&b->d_energy, sizeof(double) * Q_N * k) != hipSuccess ||
        hipHostMalloc((void **)&b->h_energy, sizeof(double) * Q_N * k, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&b->d_det, sizeof(double) * kMaxDet * k) != hipSuccess ||
        hipHostMalloc((void **)&b->h_det, sizeof(double) * kMaxDet * k, hipHostMallocDefault) != hipSuccess) {
        *out = b;
        return MDX_ERR_NOMEM;
    }
    *out = b;
    return MDX_OK;
}

mdx_status mdx_batch_set_ring(mdx_batch *b, double beadfac) {
    if (!b) return MDX_ERR_ARG;
    const int n0 = b->sys[0]->n;
    for (mdx_engine *h : b->sys)
        if (h->n != n0) {
            b->err = "path-integral beads must have the same atoms";
            return MDX_ERR_ARG;
        }
    b->ring = true;
    b->beadfac = beadfac;
    b->planned = false;
    drop_graphs(b);
    return MDX_OK;
}

mdx_status mdx_ee_energies(mdx_engine *h, const mdx_ee_params *p, int32_t nlam, const double *lambda,
                           double *e_pair, double *e_recip) {
    if (!h || !p || !lambda || nlam <= 0) return MDX_ERR_ARG;
    if (nlam > kEeMaxLam) {
        h->err = "mdx_ee_energies: at most MDX_EE_MAX_LAMBDA lambda values per call";
        return MDX_ERR_LIMIT;
    }
    if (!h->have_list) {
        h->err = "mdx_ee_energies before mdx_build_list";
        return MDX_ERR_STATE;
    }
    for (int L = 0; L < nlam; ++L)
        if (!(lambda[L] >= 0.0)) return MDX_ERR_ARG;
    hipStream_t st = h->stream;
    HIPCHK(h, hipStreamSynchronize(st));
    const int n = h->n;
    // one-time set-up: the pairs inside lambda-coupled molecules and their
    // self-term sum (DIFFUR, mdee.f:320-366)
    if (h->ee_npairs < 0) {
        std::vector<int2> pr;
        std::vector<double> q0(n);
        HIPCHK(h, hipMemcpy(q0.data(), h->d_q0, sizeof(double) * n, hipMemcpyDeviceToHost));
        const double abc = h->p.alpha / std::sqrt(kPI);
        h->ee_self0 = 0.0;
        for (int i = 0; i < n; ++i) {
            if (h->h_type_ee[h->h_moltype[i]] != 0) continue;
            h->ee_self0 += h->p.coulomb_factor * q0[i] * q0[i] * abc;
            for (int j = i + 1; j < n && h->h_mol[j] == h->h_mol[i]; ++j) pr.push_back(make_int2(i, j));
        }
        h->ee_npairs = (int)pr.size();
        CHK(upload(h, &h->d_ee_pairs, pr.data(), std::max<size_t>(1, pr.size())));
        CHK(dalloc(h, &h->d_ee_lam, 2 * kEeMaxLam));
        CHK(dalloc(h, &h->d_ee_out, 3 * kEeMaxLam));
    }
    if (h->ee_rows_cap < h->nclu_cap) {
        CHK(dalloc(h, &h->d_ee_rows, (size_t)h->nclu_cap * kEeMaxLam));
        h->ee_rows_cap = h->nclu_cap;
    }
    // the k set of the enumeration box (FURIR's 5 % window)
    std::vector<int4> ks = h->h_kidx;
    if (ks.empty() && h->p.alpha > 0.0) {
        mdx_params pk = h->p;
        for (int k = 0; k < 3; ++k) pk.box[k] = h->kbox[k];
        enumerate_k(pk, ks);
    }
    const int nkv = (int)ks.size(), nchunk = cdiv(std::max(n, 1), kFurChunk), nkb = cdiv(std::max(nkv, 1), kFurT);
    if (nkv != h->ee_nkv || std::memcmp(h->ee_kbox, h->kbox, sizeof(h->ee_kbox)) != 0) {
        CHK(upload(h, &h->d_ee_kv, ks.data(), std::max<size_t>(1, ks.size())));
        CHK(dalloc(h, &h->d_ee_spart, (size_t)nchunk * std::max(1, nkv)));
        CHK(dalloc(h, &h->d_ee_kpart, (size_t)nkb * kEeMaxLam));
        h->ee_nkv = nkv;
        std::memcpy(h->ee_kbox, h->kbox, sizeof(h->ee_kbox));
    }
    std::vector<double> lam(2 * nlam), sq(nlam);
    for (int L = 0; L < nlam; ++L) {
        sq[L] = std::pow(lambda[L], p->pw1);
        lam[2 * L] = sq[L];
        lam[2 * L + 1] = p->bee * (1.0 - std::pow(lambda[L], p->pw2));
    }
    HIPCHK(h, hipMemcpyAsync(h->d_ee_lam, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice, st));
    // pairs (DIFEN)
    EeArgs ea{};
    ea.nclu = h->nclu;
    ea.list_cap = h->list_cap;
    ea.nlam = nlam;
    ea.list = h->d_list; ea.nlist = h->d_nlist; ea.nshort = h->d_nshort; ea.clu_atom = h->d_clu_atom;
    ea.x = h->d_x; ea.y = h->d_y; ea.z = h->d_z;
    ea.box = h->box;
    ea.rc2d = h->a_pair.rc2d;
    ea.atom_mol = h->d_mol; ea.atom_moltype = h->d_moltype; ea.atom_ljtype = h->d_ljtype; ea.type_ee = h->d_type_ee;
    ea.q0 = h->d_q0;
    ea.a6_0 = h->d_a6_0; ea.b12_0 = h->d_b12_0; ea.nlj = h->nlj;
    ea.coulf = h->p.coulomb_factor;
    ea.alpha = h->p.alpha;
    ea.lam = h->d_ee_lam;
    ea.rows = h->d_ee_rows;
    if (h->nclu > 0) {
        k_ee_pairs<<<h->nclu, 64, 0, st>>>(ea);
        k_ee_colsum<<<nlam, 256, 0, st>>>(h->nclu, nlam, h->d_ee_rows, h->d_ee_out);
    } else {
        HIPCHK(h, hipMemsetAsync(h->d_ee_out, 0, sizeof(double) * nlam, st));
    }
    // reciprocal part (DIFFUR) and the intramolecular term at lambda = 1
    if (nkv > 0 && n > 0) {
        EeFurArgs fa{};
        fa.n = n; fa.nkv = nkv; fa.nchunk = nchunk; fa.nlam = nlam;
        fa.lx = h->p.box[0]; fa.ly = h->p.box[1]; fa.lz = h->p.box[2];
        fa.alp2 = h->p.alpha * h->p.alpha;
        fa.cfur = 4.0 * kPI * h->p.coulomb_factor / (h->p.box[0] * h->p.box[1] * h->p.box[2]);
        fa.x = h->d_x; fa.y = h->d_y; fa.z = h->d_z;
        fa.q = h->d_q; fa.q0 = h->d_q0;
        fa.atom_moltype = h->d_moltype; fa.type_ee = h->d_type_ee;
        fa.kv = h->d_ee_kv; fa.spart = h->d_ee_spart; fa.lam = h->d_ee_lam; fa.kpart = h->d_ee_kpart;
        k_ee_sk<<<dim3(nkb, nchunk), kFurT, 0, st>>>(fa);
        k_ee_sred<<<nkb, kFurT, 0, st>>>(fa);
        k_ee_colsum<<<nlam, 256, 0, st>>>(nkb, nlam, h->d_ee_kpart, h->d_ee_out + kEeMaxLam);
    } else {
        HIPCHK(h, hipMemsetAsync(h->d_ee_out + kEeMaxLam, 0, sizeof(double) * nlam, st));
    }
    k_ee_intra<<<1, 256, 0, st>>>(h->ee_npairs, h->d_ee_pairs, h->d_x, h->d_y, h->d_z, h->d_q0, h->box,
                                  h->p.coulomb_factor, h->p.alpha, h->d_ee_out + 2 * kEeMaxLam);
    std::vector<double> out(3 * kEeMaxLam);
    HIPCHK(h, hipMemcpyAsync(out.data(), h->d_ee_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    HIPCHK(h, hipGetLastError());
    const double intra0 = out[2 * kEeMaxLam];
    for (int L = 0; L < nlam; ++L) {
        if (e_pair) e_pair[L] = out[L];
        const double s2 = sq[L] * sq[L];
        if (e_recip) e_recip[L] = out[kEeMaxLam + L] - s2 * h->ee_self0 + s2 * intra0;
    }
    return MDX_OK;
}

mdx_status mdx_get_terms(mdx_engine *h, double *br, double *be, double *ar, double *ae, double *tr, double *te) {
    if (!h) return MDX_ERR_ARG;
    std::vector<double> v(2 * (size_t)std::max(1, h->ntm));
    HIPCHK(h, hipStreamSynchronize(h->self->stream));
    HIPCHK(h, hipMemcpy(v.data(), h->d_terms, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
    auto out = [&](double *val, double *en, int m0, int cnt) {
        for (int m = 0; m < cnt; ++m) {
            if (val) val[m] = v[2 * (size_t)(m0 + m)];
            if (en) en[m] = v[2 * (size_t)(m0 + m) + 1];
        }
    };
    out(br, be, 0, h->nbtm);
    out(ar, ae, h->nbtm, h->natm);
    out(tr, te, h->nbtm + h->natm, h->nttm);
    return MDX_OK;
}

mdx_status mdx_batch_use_graph(mdx_batch *b, int32_t on) {
    if (!b) return MDX_ERR_ARG;
    b->use_graph = on != 0;
    return MDX_OK;
}

mdx_status mdx_batch_compute(mdx_batch *b, int32_t group, int32_t want, mdx_energies *e) {
    if (!b) return MDX_ERR_ARG;
    const bool detail = (want & MDX_WANT_DETAIL) != 0, terms = (want & MDX_WANT_TERMS) != 0;
    CHK(batch_compute(b, group, detail, terms));
    if (e)
        for (size_t s = 0; s < b->sys.size(); ++s) fill_energies(b, (int)s, group, detail, e + s);
    return MDX_OK;
}

const char *mdx_batch_last_error(const mdx_batch *b) { return b ? b->err.c_str() : "null batch"; }

mdx_status mdx_batch_destroy(mdx_batch *b) {
    if (!b) return MDX_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graphs(b);
    kplan_free(b->pack); kplan_free(b->com); kplan_free(b->pairs); kplan_free(b->pairs64);
    kplan_free(b->special);
    kplan_free(b->rprep); kplan_free(b->sk); kplan_free(b->skred); kplan_free(b->rforce);
    kplan_free(b->et); kplan_free(b->bonds); kplan_free(b->spc); kplan_free(b->angles); kplan_free(b->tors); kplan_free(b->terms);
    kplan_free(b->springs); kplan_free(b->comb); kplan_free(b->red); kplan_free(b->redp); kplan_free(b->dred);
    kplan_free(b->md_kd); kplan_free(b->md_sh); kplan_free(b->md_ke); kplan_free(b->md_th);
    kplan_free(b->md_gc); kplan_free(b->md_tr);
    if (b->d_trace) hipFree(b->d_trace);
    if (b->d_xi) hipFree(b->d_xi);
    if (b->d_vptr) hipFree(b->d_vptr);
    if (b->d_energy) hipFree(b->d_energy);
    if (b->h_energy) hipHostFree(b->h_energy);
    if (b->d_det) hipFree(b->d_det);
    if (b->h_det) hipHostFree(b->h_det);
    for (int e = 0; e < 12; ++e)
        if (b->ev[e]) hipEventDestroy(b->ev[e]);
    for (int k = 0; k < 2; ++k) {
        if (b->join_ev[k]) hipEventDestroy(b->join_ev[k]);
        if (b->side[k]) hipStreamDestroy(b->side[k]);
    }
    if (b->fork_ev) hipEventDestroy(b->fork_ev);
    if (b->own_stream && b->stream) hipStreamDestroy(b->stream);
    delete b;
    return MDX_OK;
}

}  // extern "C"

#include "serve.hip"
