// Real-space pair forces (replaces LOCALF forces.f:813-998 and FORCES
// forces.f:1014-1202).
//
// pairs_body: one wavefront per i-cluster of 8 atoms; each iteration takes
// one j-cluster of its list and lane (i_slot * 8 + j_slot) evaluates one
// atom pair in FP32: Lennard-Jones plus the Abramowitz-Stegun erfc-screened
// Ewald real-space term (the reference's polynomial, prcm.h:27-29, with the
// exact Gaussian in the force, forces.f:1146-1153).  Distance vectors are
// formed from anchor-relative FP32 coordinates plus the cluster pair's
// image shift, so no FP32 absolute coordinate enters a difference.
// The list is a half list (each atom pair once): the i atoms' forces are
// summed by the wavefront that owns the i-cluster, in a fixed order, with
// FP32 partial sums over 16 j-clusters flushed into FP64; the j atoms'
// reactions (Newton's third law) are summed over the 8 i slots in a fixed
// butterfly and added into 64-bit fixed-point accumulators, whose integer
// sums do not depend on the order of arrival.  Energies and virials go to
// per-wavefront FP64 partials.  The
// electrostatic energy is accumulated per unit q_i C with the exact q_j
// (FP32 high part plus low part) and multiplied by the FP64 q_i C at the
// end, so no charge rounding biases the cancelling sums.  Pairs keep the
// short/long class of the list build: short pairs feed the fast forces H,
// long pairs the slow forces G.
//
// special_body: the 1-4 and lambda-coupled intramolecular pairs in FP64
// with the reference's exact expressions (direct Coulomb, 1-4 scaling).
#include "mdx_internal.h"

#include <type_traits>

namespace mdx {

// Anchor-relative packed coordinates of every cluster slot.
__device__ inline void pack_body(const PackArgs &a, int b) {
    const int s = b * 256 + threadIdx.x;
    if (s >= a.nclu * kClu) return;
    const int at = a.clu_atom[s];
    const int c = s / kClu;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (at >= 0) {
        double d[3] = {a.x[at] - a.anchor[c * 3], a.y[at] - a.anchor[c * 3 + 1], a.z[at] - a.anchor[c * 3 + 2]};
        for (int k = 0; k < 3; ++k) d[k] -= a.box.l[k] * rint(d[k] / a.box.l[k]);
        o = make_float4((float)d[0], (float)d[1], (float)d[2], (float)a.q[at]);
    }
    a.xq[s] = o;
}

__global__ __launch_bounds__(256) void k_pack(const PackArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    pack_body(args[m.x], m.y);
}

// Molecular centres of mass (GETCOM, service.f:471-500) for the molecular
// virial.  A wavefront takes one group of consecutive molecules (at most
// 64, at most kComAtoms atoms unless one molecule is larger): their atoms
// pass through LDS in coalesced stages, and each lane adds its molecule's
// atoms in atom order.
__device__ inline void com_body(const PackArgs &a, int blk, double *sh) {
    if (a.user_com) return;
    const int lane = threadIdx.x;
    const int m0 = a.comgrp[blk], m1 = a.comgrp[blk + 1];
    const int m = m0 + lane;
    const int A0 = a.mstart[m0], A1 = a.mstart[m1];
    const int b0 = m < m1 ? a.mstart[m] : A1, b1 = m < m1 ? a.mstart[m + 1] : A1;
    double *sm_ = sh, *sx = sh + kComAtoms, *sy = sh + 2 * kComAtoms, *sz = sh + 3 * kComAtoms;
    double cx = 0.0, cy = 0.0, cz = 0.0, sm = 0.0;
    for (int c0 = A0; c0 < A1; c0 += kComAtoms) {
        const int c1 = min(c0 + kComAtoms, A1);
        for (int i = c0 + lane; i < c1; i += 64) {
            sm_[i - c0] = a.mass[i]; sx[i - c0] = a.x[i]; sy[i - c0] = a.y[i]; sz[i - c0] = a.z[i];
        }
        __syncthreads();
        for (int i = max(b0, c0); i < min(b1, c1); ++i) {
            const double mi = sm_[i - c0];
            cx = cx + mi * sx[i - c0];
            cy = cy + mi * sy[i - c0];
            cz = cz + mi * sz[i - c0];
            sm = sm + mi;
        }
        __syncthreads();
    }
    if (m < m1) {
        a.cx[m] = cx / sm;
        a.cy[m] = cy / sm;
        a.cz[m] = cz / sm;
    }
}

__global__ __launch_bounds__(64) void k_com(const PackArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double sh[4 * kComAtoms];
    const int2 m = map[blockIdx.x];
    com_body(args[m.x], m.y, sh);
}

// Reduce-scatter steps on the gfx950 lane-swap instructions.
// swap_sum32(x, y): lanes 0-31 get x + x(lane + 32), lanes 32-63 get
// y(lane - 32) + y; v_permlane32_swap exchanges x's upper half with y's
// lower half.
__device__ inline float swap_sum32(float x, float y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// swap_sum16(x, y): even rows of 16 lanes get x + x(lane + 16), odd rows
// y(lane - 16) + y (v_permlane16_swap: x's odd rows with y's even rows).
__device__ inline float swap_sum16(float x, float y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// a * b rounded on its own: never contracted into a later add
__device__ inline float mul_unfused(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}

// x + x(lane ^ 8): the two halves of each row of 16 lanes added (DPP
// row_ror:8; every lane reads a lane of its own row, so the move has no
// old value to keep and folds into the add as its DPP operand).
__device__ inline float add_xor8(float x) {
    return x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x128, 0xf, 0xf, false));
}

// j reactions of the half list leave the pair kernel as a record per list
// entry at the entry's record position (PairArgs::rec, jdest): plain
// stores, marked non-temporal (streamed to memory, not kept in the caches,
// where the next call's stores to the same lines would find them), no
// atomics; k_jgather sums them per j atom.  The butterflies below leave
// the 8 j slots' sums in lanes 16 row + j_slot of each row of 16 lanes,
// written to record floats 8 (row) + j_slot, and 32 + 8 (row / 2) + j_slot
// for the second value of a mixed entry:
//   long-class entry:  rows 0, 1, 2 = long x, z, y             (floats 0-23)
//   mixed entry:       first value rows 0-3 = long x, z, y, short x;
//                      second value rows 0, 2 = short y, z     (floats 0-47)
constexpr unsigned long long kPatLong = 0x000000ff00ff00ffull;
constexpr unsigned long long kPatMix0 = 0x00ff00ff00ff00ffull;
constexpr unsigned long long kPatMix1 = 0x000000ff000000ffull;
constexpr double kMagic = 6755399441055744.0;              // 1.5 * 2^52
constexpr int kPairBlock = 16;      // list entries per FP32 partial sum

// FP32 partial sums of one block of list entries (per lane).
template <bool MIXED>
struct PairAcc {
    float fl[3] = {0.f, 0.f, 0.f}, vl[3] = {0.f, 0.f, 0.f};
    float el = 0.f, el_lo = 0.f, lj = 0.f, vr = 0.f;          // long class
    // long-class entries: 12 E_LJ as the sums of 6 A6 / r^6 and of r F_LJ
    // (= vr except for soft-core pairs, which keep their own w6 sum)
    float a6 = 0.f, w6 = 0.f;
    float fs[MIXED ? 3 : 1] = {}, vs[MIXED ? 3 : 1] = {};
    float es = 0.f, ljs = 0.f, vrs = 0.f;                      // short class
};

// Detail calls (MDX_WANT_DETAIL) also split the bare Coulomb energy
// q_i q_j C / (r + RAE) and the LJ energy by molecule-type pair (POTE1,
// POTL1) or, for pairs inside one molecule, by molecule type (PES141,
// PSR141), as LOCALF/FORCES do (forces.f:945-974, 1155-1183).  Each lane
// keeps its own FP64 buckets in LDS (bucket = type of the j molecule, or
// the intramolecular bucket), so the sums stay in a fixed order.
// Detail row of one i-cluster from the per-lane buckets: bucket b of lane
// L holds the Coulomb (2 b) and LJ (2 b + 1) energies of lane L's pairs
// whose j molecule has type b, or (b = nmt) lies in the i molecule; each
// output column sums its (lane, bucket) entries in lane order.  The pair
// type index is symmetric, so each pair counts once in its bucket.
__device__ inline void detail_out(const PairArgs &a, int ci, const double *s_det, const int *s_ti) {
    for (int o = threadIdx.x; o < a.ndet; o += 64) {
        double sum = 0.0;
        for (int L = 0; L < 64; ++L) {
            const int ti = s_ti[L >> 3];
            if (o < 2 * a.ntp) {
                const int k = o < a.ntp ? 0 : 1, mtr = o < a.ntp ? o : o - a.ntp;
                for (int b = 0; b < a.nmt; ++b)
                    if (a.type_pair[ti * a.nmt + b] == mtr) sum += s_det[(b * 2 + k) * 64 + L];
            } else {
                const int r = o - 2 * a.ntp;
                const int k = r < a.nmt ? 0 : 1, t = r < a.nmt ? r : r - a.nmt;
                if (ti == t) sum += s_det[(a.nmt * 2 + k) * 64 + L];
            }
        }
        a.dpart[(size_t)ci * a.ndet + o] = sum;
    }
}

// (inlined into every kernel instantiation, which then gets its own
// register budget)
template <bool SOFT, bool DETAIL>
__device__ __attribute__((always_inline)) inline void pairs_body(const PairArgs &a, int ci, float2 *s_lj,
                                                                 double *s_acc, int group) {
    const int lane = threadIdx.x;
    const int nlj2 = a.nlj * a.nlj;
    for (int t = lane; t < nlj2; t += 64) s_lj[t] = ldv(gld(a.lj) + t);
    double *s_det = reinterpret_cast<double *>(s_lj + nlj2);   // [(nmt + 1) * 2][64]
    int *s_ti = reinterpret_cast<int *>(s_det + (a.nmt + 1) * 2 * 64);
    if (DETAIL) {
        for (int b = 0; b < (a.nmt + 1) * 2; ++b) s_det[b * 64 + lane] = 0.0;
        if (ci >= a.nclu) {
            for (int o = lane; o < a.ndet; o += 64) a.dpart[(size_t)ci * a.ndet + o] = 0.0;
            return;
        }
    }
    __syncthreads();
    if (ci >= a.nclu) return;

    const int il = lane >> 3;
    const int jl = lane & 7;
    const float4 pi = ldv(gld(a.xq) + ci * kClu + il);
    const int infoi = gld(a.aux)[ci * kClu + il].x;
    const float2 *ljrow = s_lj + (infoi & 0xffff) * a.nlj;   // (6 A6, 12 B12) against type j
    const float A1 = (float)kA1, A2 = (float)kA2, A3 = (float)kA3, A4 = (float)kA4, A5 = (float)kA5;
    const float alpha = a.alpha;
    const float b1a = (float)kB1 * alpha;
    const float nal2 = -alpha * alpha * 1.44269504088896341f;   // -alpha^2 log2(e)
    const float talpha = (float)kTOTPI * alpha;
    const float qic = pi.w * a.coulf_f;
    int moli = -1;
    if (DETAIL) {
        const int ai0 = a.clu_atom[ci * kClu + il];
        moli = ai0 >= 0 ? a.atom_mol[ai0] : -1;
        if (jl == 0) s_ti[il] = ai0 >= 0 ? a.atom_moltype[ai0] : 0;
    }

    double Fs[3] = {0, 0, 0}, Fl[3] = {0, 0, 0};
    // the FP64 energy and virial sums per lane live in LDS (touched once per
    // block of entries; registers stay free for the pair loop): virial
    // short xyz, long xyz, then energies el_s, el_l, lj_s, lj_l, vlj_s, vlj_l
    double *const sE = s_acc + 6 * 64 + lane;
    for (int q = 0; q < 12; ++q) s_acc[q * 64 + lane] = 0.0;
    int cnt_s = 0, cnt_l = 0;   // wave totals (ballot counts)
    const int ns = sld(a.nshort)[ci];
    const int n = sld(a.nlist)[ci];
    const MDX_CONST PairEntry *L = sld(a.list) + (size_t)ci * a.list_cap;
    const float bl[3] = {a.lx, a.ly, a.lz}, bi[3] = {a.ilx, a.ily, a.ilz};
    const float rlo = a.rlo, rhi = a.rhi, raex = a.raex;
    // j-cluster data through buffer descriptors: lane offset in a VGPR,
    // cluster offset in an SGPR
    const __amdgpu_buffer_rsrc_t rs_xq =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.xq, (short)0, a.nclu * kClu * 16, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs_aux =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.aux, (short)0, a.nclu * kClu * 8, 0x00020000);
    const int voff_xq = jl * 16, voff_aux = jl * 8;
    auto ld_xq = [&](int cj) {
        return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs_xq, voff_xq, cj * (kClu * 16), 0));
    };
    auto ld_aux = [&](int cj) {
        return __builtin_bit_cast(int2, __builtin_amdgcn_raw_buffer_load_b64(rs_aux, voff_aux, cj * (kClu * 8), 0));
    };
    auto jclu = [](const PairEntry &e) { return (int)((unsigned)e.cj & ~kEntryWrap); };

    // j-reaction records: record base address in SGPRs, lane float in a VGPR
    const int fr0 = (lane >> 4) * 8 + jl, fr1 = 32 + (lane >> 5) * 8 + jl;
    const MDX_CONST unsigned *jdest = sld(a.jdest) + (size_t)ci * a.list_cap;
    auto put = [&](float v, unsigned long long pat, int f, unsigned dest) {
        if (__builtin_amdgcn_inverse_ballot_w64(pat))
            __builtin_nontemporal_store(v, gst(a.rec + (size_t)dest * kRecUnit) + f);
    };
    // FP64 cutoff test of this lane's pair with the reference arithmetic
    auto exact_in = [&](int cj) {
#pragma clang fp contract(off)
        const int ai = gld(a.clu_atom)[ci * kClu + il], aj = gld(a.clu_atom)[cj * kClu + jl];
        double dx = gld(a.x)[ai] - gld(a.x)[aj], dy = gld(a.y)[ai] - gld(a.y)[aj];
        double dz = gld(a.z)[ai] - gld(a.z)[aj];
        min_image(dx, dy, dz, a.box);
        double rr2 = dx * dx + dy * dy;
        rr2 = rr2 + dz * dz;
        return rr2 <= a.rc2d;
    };

    // One list entry.  MIXED: an entry of the front segment (it holds a
    // short-class pair, so both classes are split by exact 0/1 weights);
    // otherwise every pair is long-class and only the long sums move.
    // Branch-free body: a lane without a pair inside the cutoff takes
    // rinv = 0, which makes every contribution exactly zero.
    auto entry = [&](auto mixed, PairAcc<decltype(mixed)::value> &s, const PairEntry &e, float4 pj, int2 auxj,
                     unsigned slot) {   // slot: the entry's record position (units)
        constexpr bool MIXED = decltype(mixed)::value;
        // uses of the j data stay here, not next to their prefetch
        __builtin_amdgcn_sched_barrier(0);
        unsigned long long m = e.mask;
        if (group == 1) m &= e.mshort;
        else if (group == 2) m &= ~e.mshort;
        float dx = (pi.x - pj.x) + e.sx;
        float dy = (pi.y - pj.y) + e.sy;
        float dz = (pi.z - pj.z) + e.sz;
        // entries whose clusters reach near half the box: minimum image
        // per pair (the flag is uniform over the wavefront)
        if (e.cj < 0) {
            dx = fmaf(-bl[0], rintf(dx * bi[0]), dx);
            dy = fmaf(-bl[1], rintf(dy * bi[1]), dy);
            dz = fmaf(-bl[2], rintf(dz * bi[2]), dz);
        }
        const float dx2 = dx * dx;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx2));
        // lanes in the list and inside the cutoff, as a wave mask; pairs
        // within the FP32 error band of the cutoff take the reference's
        // FP64 test (no band for reference lists: rlo = rhi)
        unsigned long long inm = __ballot(r2 <= rlo) & m;
        const unsigned long long amb = __ballot(r2 <= rhi) & m & ~inm;
        if (amb != 0ull) {
            const bool mine = (amb >> lane) & 1ull;
            inm |= __ballot(mine && exact_in(jclu(e)));
        }
        const bool in = __builtin_amdgcn_inverse_ballot_w64(inm);
        const float rinv = in ? __builtin_amdgcn_rsqf(r2) : 0.f;
        const float r = r2 * rinv;
        float rei = rinv;
        bool softp = false;      // a soft-core pair (lambda-coupled types)
        if (SOFT) {
            const int eej = auxj.x >> 16, eei = infoi >> 16;
            softp = in && (eei & eej) == 0;
            rei = softp ? __builtin_amdgcn_rcpf(r + raex) : rinv;
        }
        const float t = __builtin_amdgcn_rcpf(fmaf(b1a, r, 1.0f));
        const float g = __builtin_amdgcn_exp2f(nal2 * r2);
        const float pt = ((((A5 * t + A4) * t + A3) * t + A2) * t + A1) * t;
        const float phi = pt * (g * rei);                   // erfc(alpha r) / r
        const float2 lj = ljrow[auxj.x & 0xffff];
        const float r2i = rei * rei;
        const float r6i = r2i * r2i * r2i;
        // r F_LJ; a rounded product (not fused into a later sum), so both
        // soft-core instantiations form the same sums for ordinary pairs
        const float w6 = mul_unfused(fmaf(lj.y, r6i, lj.x), r6i);
        const float e12 = fmaf(lj.x, r6i, w6);              // 12 E_LJ
        const float qq = qic * pj.w;
        const float t2 = fmaf(qq, fmaf(talpha, g, phi), w6);
        const float fsc = (SOFT ? rinv * rei : r2i) * t2;   // |F| / r
        // every other pair takes the same operations in both instantiations
        const float vr = softp ? w6 * (rei * r) : w6;
        const float qlo = __int_as_float(auxj.y);
        if (DETAIL && in) {
            const int aj = a.clu_atom[jclu(e) * kClu + jl];
            const int b = a.atom_mol[aj] == moli ? a.nmt : a.atom_moltype[aj];
            s_det[(b * 2) * 64 + lane] += (double)(pj.w * rei) + (double)(qlo * rei);
            s_det[(b * 2 + 1) * 64 + lane] += (double)e12 * (1.0 / 12.0);
        }
        if (MIXED) {
            const bool isshort = __builtin_amdgcn_inverse_ballot_w64(e.mshort);
            const float ws = isshort ? 1.f : 0.f, wl = 1.f - ws;
            // each class's |F| / r is fsc or exactly 0 (a product with its
            // 0/1 weight), and its reactions on j (minus the force on i)
            // are formed from it, so the other class gets exact zeros; a
            // difference such as fsc - ws fsc could keep the rounding error
            // of a product that contraction fused into it
            const float fss = ws * fsc, fsl = wl * fsc;
            const float sx = -fss * dx, sy = -fss * dy, sz = -fss * dz;
            const float lx = -fsl * dx, ly = -fsl * dy, lz = -fsl * dz;
            s.fs[0] = fmaf(fss, dx, s.fs[0]); s.fs[1] = fmaf(fss, dy, s.fs[1]); s.fs[2] = fmaf(fss, dz, s.fs[2]);
            s.fl[0] = fmaf(fsl, dx, s.fl[0]); s.fl[1] = fmaf(fsl, dy, s.fl[1]); s.fl[2] = fmaf(fsl, dz, s.fl[2]);
            // virial f . d per axis from the reactions (-sx dx = fss dx^2)
            s.vs[0] = fmaf(-sx, dx, s.vs[0]); s.vs[1] = fmaf(-sy, dy, s.vs[1]); s.vs[2] = fmaf(-sz, dz, s.vs[2]);
            s.vl[0] = fmaf(-lx, dx, s.vl[0]); s.vl[1] = fmaf(-ly, dy, s.vl[1]); s.vl[2] = fmaf(-lz, dz, s.vl[2]);
            const float ees = fmaf(qlo, phi, pj.w * phi);
            s.es = fmaf(ws, ees, s.es); s.ljs = fmaf(ws, e12, s.ljs); s.vrs = fmaf(ws, vr, s.vrs);
            s.el = fmaf(wl, ees, s.el); s.lj = fmaf(wl, e12, s.lj); s.vr = fmaf(wl, vr, s.vr);
            cnt_s += __popcll(inm & e.mshort);
            cnt_l += __popcll(inm & ~e.mshort);
            const float d0 = add_xor8(swap_sum16(swap_sum32(lx, ly), swap_sum32(lz, sx)));
            const float d1 = add_xor8(swap_sum16(swap_sum32(sy, sz), 0.f));
            put(d0, kPatMix0, fr0, slot);
            put(d1, kPatMix1, fr1, slot);
        } else {
            // reactions on j (minus the force on i)
            const float nx = -fsc * dx, ny = -fsc * dy, nz = -fsc * dz;
            s.fl[0] = fmaf(fsc, dx, s.fl[0]); s.fl[1] = fmaf(fsc, dy, s.fl[1]); s.fl[2] = fmaf(fsc, dz, s.fl[2]);
            // (the reactions go into destructive lane swaps below; squares
            // of d here need no copies of them)
            s.vl[0] = fmaf(fsc, dx2, s.vl[0]); s.vl[1] = fmaf(fsc, dy * dy, s.vl[1]);
            s.vl[2] = fmaf(fsc, dz * dz, s.vl[2]);
            s.el = fmaf(pj.w, phi, s.el); s.el_lo = fmaf(qlo, phi, s.el_lo);
            s.a6 = fmaf(lj.x, r6i, s.a6);
            if (SOFT) s.w6 += w6;
            s.vr += vr;
            cnt_l += __popcll(inm);
            const float d = add_xor8(swap_sum16(swap_sum32(nx, ny), swap_sum32(nz, 0.f)));
            put(d, kPatLong, fr0, slot);
        }
    };

    // One segment of the list: entries P[0], P[step], ..., cnt of them,
    // two per iteration (the j data of each entry is loaded one entry
    // ahead, the entries themselves two ahead, in fixed registers)
    auto segment = [&](auto mixed, int slot0, int step, int cnt) {
        const MDX_CONST PairEntry *P = L + slot0;
        constexpr bool MIXED = decltype(mixed)::value;
        if (cnt <= 0) return;
        auto at = [&](int k) { return ldv(P + (ptrdiff_t)step * min(k, cnt - 1)); };
        PairEntry eA = at(0), eB = at(1);
        float4 pjA = ld_xq(jclu(eA));
        int2 auxA = ld_aux(jclu(eA));
        for (int k0 = 0; k0 < cnt; k0 += kPairBlock) {
            PairAcc<MIXED> s;
            const int k1 = min(k0 + kPairBlock, cnt);
            for (int k = k0; k < k1; k += 2) {
                const float4 pjB = ld_xq(jclu(eB));
                const int2 auxB = ld_aux(jclu(eB));
                const PairEntry eC = at(k + 2);
                entry(mixed, s, eA, pjA, auxA, ldv(jdest + slot0 + step * k));
                pjA = ld_xq(jclu(eC));
                auxA = ld_aux(jclu(eC));
                const PairEntry eD = at(k + 3);
                if (k + 1 < k1) entry(mixed, s, eB, pjB, auxB, ldv(jdest + slot0 + step * (k + 1)));
                eA = eC;
                eB = eD;
            }
            for (int q = 0; q < 3; ++q) Fl[q] += (double)s.fl[q];
            sE[64] += (double)s.el + (double)s.el_lo;
            sE[192] += ((double)s.lj + (double)s.a6 + (double)(MIXED ? 0.f : SOFT ? s.w6 : s.vr)) * (1.0 / 12.0);
            sE[320] += (double)s.vr;
            for (int q = 0; q < 3; ++q) s_acc[(3 + q) * 64 + lane] += (double)s.vl[q];
            if (MIXED) {
                for (int q = 0; q < 3; ++q) Fs[q] += (double)s.fs[q];
                sE[0] += (double)s.es;
                sE[128] += (double)s.ljs * (1.0 / 12.0);
                sE[256] += (double)s.vrs;
                for (int q = 0; q < 3; ++q) s_acc[q * 64 + lane] += (double)s.vs[q];
            }
        }
    };
    // the front segment holds the entries with a short-class pair; a
    // fast-class call reads only these; the others sit at the back of the
    // cluster's list space, last first
    segment(std::true_type{}, 0, 1, ns);
    if (group != 1) segment(std::false_type{}, a.list_cap - 1, -1, n - ns);

    // i forces: sum over the 8 lanes of one i slot
    for (int off = 1; off < 8; off <<= 1)
        for (int q = 0; q < 3; ++q) {
            Fs[q] += __shfl_xor(Fs[q], off);
            Fl[q] += __shfl_xor(Fl[q], off);
        }
    const int ai = gld(a.clu_atom)[ci * kClu + il];
    const double qc = ai >= 0 ? gld(a.q)[ai] * a.coulf : 0.0;
    double Vs[3], Vl[3];
    for (int q = 0; q < 3; ++q) { Vs[q] = s_acc[q * 64 + lane]; Vl[q] = s_acc[(3 + q) * 64 + lane]; }
    double E[6];
    for (int q = 0; q < 6; ++q) E[q] = sE[q * 64];
    E[0] *= qc;
    E[1] *= qc;
    if (jl == 0 && ai >= 0) {
        if (a.fs_x) { gst(a.fs_x)[ai] = Fs[0]; gst(a.fs_y)[ai] = Fs[1]; gst(a.fs_z)[ai] = Fs[2]; }
        if (a.fl_x) { gst(a.fl_x)[ai] = Fl[0]; gst(a.fl_y)[ai] = Fl[1]; gst(a.fl_z)[ai] = Fl[2]; }
    }
    // energies and virials: wavefront sums (every pair once)
    double q[14] = {E[0], E[1], E[2], E[3], E[4], E[5], Vs[0], Vs[1], Vs[2], Vl[0], Vl[1], Vl[2],
                    lane == 0 ? (double)cnt_s : 0.0, lane == 0 ? (double)cnt_l : 0.0};
    for (int off = 1; off < 64; off <<= 1)
        for (int k = 0; k < 14; ++k) q[k] += __shfl_xor(q[k], off);
    if (lane == 0) {
        MDX_GLOBAL double *o = gst(a.epart) + (size_t)ci * Q_N;
        for (int k = 0; k < Q_N; ++k) o[k] = 0.0;
        o[Q_PELS2] = q[0];
        o[Q_PELS1] = q[1];
        o[Q_PE2] = q[2];
        o[Q_PE1] = q[3];
        o[Q_VIR2] = q[4];
        o[Q_VIR1] = q[5];
        o[Q_VIRFX] = q[6]; o[Q_VIRFY] = q[7]; o[Q_VIRFZ] = q[8];
        o[Q_VIRX] = q[9]; o[Q_VIRY] = q[10]; o[Q_VIRZ] = q[11];
        o[Q_NPS] = q[12];
        o[Q_NPL] = q[13];
    }
    if (DETAIL) {
        // bare Coulomb buckets per unit q_i C -> energies
        for (int b = 0; b <= a.nmt; ++b) s_det[(b * 2) * 64 + lane] *= qc;
        __syncthreads();
        detail_out(a, ci, s_det, s_ti);
    }
}

// WPE wavefronts per SIMD: the plain kernel comes at six (79 VGPRs, no
// spill: 4-5 % faster than five at 96,174 atoms) and at five, and the
// autotuner picks one per size class; the soft-core and detail
// instantiations hold more live values and run at five.
template <bool SOFT, bool DETAIL, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_pairs(const PairArgs *__restrict__ args, const int2 *__restrict__ map, int group) {
    extern __shared__ float2 s_lj[];      // nlj x nlj (6 A6, 12 B12), then detail buckets
    __shared__ double s_acc[12 * 64];     // per-lane FP64 virial and energy sums
    const int2 m = map[blockIdx.x];
    pairs_body<SOFT, DETAIL>(args[m.x], m.y, s_lj, s_acc, group);
}

template __global__ void k_pairs<false, false, 6>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<false, false, 5>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<true, false, 5>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<false, true, 5>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<true, true, 5>(const PairArgs *, const int2 *, int);

// FP64 precision mode: the pairs of the same list, each in FP64 from the
// caller's coordinates as FORCES and LOCALF form it (forces.f:1078-1183,
// 870-990): the one-shift minimum image of the atom pair, r = sqrt(r^2),
// the A&S polynomial times exp(-(alpha r)^2), the soft-core shift RAE of
// lambda-coupled pairs in 1/r but not in the erfc argument, LJ only where
// B12 != 0.  Lane (i_slot * 8 + j_slot) as in pairs_body, one list entry
// at a time (front segment, then the long-class entries); the i forces
// are summed per lane in FP64 in list order; the j reactions are summed
// over the 8 i slots by a butterfly of FP64 shuffles and added into the j
// slots' 64-bit fixed-point words (PairArgs::jf, cleared before the call)
// with integer atomics, whose totals do not depend on the order of
// arrival.  k_jgather does not run in this mode.
template <bool SOFT, bool DETAIL>
__device__ inline void pairs64_body(const PairArgs &a, int ci, double *s_det, int group) {
    const int lane = threadIdx.x;
    int *s_ti = reinterpret_cast<int *>(s_det + (a.nmt + 1) * 2 * 64);
    if (DETAIL) {
        for (int b = 0; b < (a.nmt + 1) * 2; ++b) s_det[b * 64 + lane] = 0.0;
        if (ci >= a.nclu) {
            for (int o = lane; o < a.ndet; o += 64) a.dpart[(size_t)ci * a.ndet + o] = 0.0;
            return;
        }
    }
    __syncthreads();
    if (ci >= a.nclu) return;
    const int il = lane >> 3, jl = lane & 7;
    const int ai = a.clu_atom[ci * kClu + il];
    const int infoi = a.aux[ci * kClu + il].x;
    const double xi = ai >= 0 ? a.x[ai] : 0.0, yi = ai >= 0 ? a.y[ai] : 0.0, zi = ai >= 0 ? a.z[ai] : 0.0;
    const double qic = ai >= 0 ? a.q[ai] * a.coulf : 0.0;
    const double2 *ljrow = a.lj64 + (size_t)(infoi & 0xffff) * a.nlj;
    int moli = -1;
    if (DETAIL) {
        moli = ai >= 0 ? a.atom_mol[ai] : -1;
        if (jl == 0) s_ti[il] = ai >= 0 ? a.atom_moltype[ai] : 0;
    }
    // per lane: forces and virial projections (short, long), energies
    // Coulomb short/long, LJ short/long, LJ virial short/long
    double Fs[3] = {0, 0, 0}, Fl[3] = {0, 0, 0}, Vs[3] = {0, 0, 0}, Vl[3] = {0, 0, 0};
    double E[6] = {0, 0, 0, 0, 0, 0};
    int cnt_s = 0, cnt_l = 0;
    const int ns = a.nshort[ci], n = a.nlist[ci];
    const PairEntry *L = a.list + (size_t)ci * a.list_cap;
    const int nk = group == 1 ? ns : n;
    for (int k = 0; k < nk; ++k) {
        const bool front = k < ns;
        const PairEntry e = L[front ? k : a.list_cap - 1 - (k - ns)];
        const int cj = (int)((unsigned)e.cj & ~kEntryWrap);
        unsigned long long m = e.mask;
        if (group == 1) m &= e.mshort;
        else if (group == 2) m &= ~e.mshort;
        const int sj = cj * kClu + jl;
        const bool isshort = (e.mshort >> lane) & 1ull;
        double d[3] = {0.0, 0.0, 0.0}, force = 0.0;
        if ((m >> lane) & 1ull) {
#pragma clang fp contract(off)
            const int aj = a.clu_atom[sj];
            d[0] = xi - a.x[aj];
            d[1] = yi - a.y[aj];
            d[2] = zi - a.z[aj];
            min_image(d[0], d[1], d[2], a.box);
            double rr = d[0] * d[0] + d[1] * d[1];
            rr = rr + d[2] * d[2];
            if (rr <= a.rc2d) {
                const int infoj = a.aux[sj].x;
                const double r1 = sqrt(rr), r1i = 1.0 / r1;
                const double rae = (SOFT && ((infoi >> 16) & (infoj >> 16)) == 0) ? a.raexd : 0.0;
                const double r1ei = 1.0 / (r1 + rae);
                const double ar = a.alphad * r1;
                const double t = 1.0 / (1.0 + kB1 * ar);
                const double ex = exp(-(ar * ar));
                const double erfct = ((((kA5 * t + kA4) * t + kA3) * t + kA2) * t + kA1) * t * ex;
                const double ee1 = qic * a.q[aj] * r1ei;          // bare Coulomb
                const double ees = ee1 * erfct;
                force = ee1 * (kTOTPI * a.alphad * ex + erfct * r1ei) * r1i;
                const double2 lj = ljrow[infoj & 0xffff];        // (A6, B12)
                double esr = 0.0, vr = 0.0;
                if (lj.y != 0.0) {
                    const double r2i = r1ei * r1ei, r6i = r2i * r2i * r2i;
                    esr = (lj.x + lj.y * r6i) * r6i;
                    const double fsr = (6.0 * lj.x + 12.0 * lj.y * r6i) * r6i * r1ei;
                    force = force + fsr * r1i;
                    vr = fsr * r1;
                }
                const int c = isshort ? 0 : 1;
                E[c] += ees;
                E[2 + c] += esr;
                E[4 + c] += vr;
                double *F = isshort ? Fs : Fl, *V = isshort ? Vs : Vl;
                for (int q = 0; q < 3; ++q) {
                    F[q] += force * d[q];
                    V[q] += force * (d[q] * d[q]);
                }
                if (isshort) ++cnt_s;
                else ++cnt_l;
                if (DETAIL) {
                    const int b = a.atom_mol[aj] == moli ? a.nmt : a.atom_moltype[aj];
                    s_det[(b * 2) * 64 + lane] += ee1;
                    s_det[(b * 2 + 1) * 64 + lane] += esr;
                }
            }
        }
        // reactions on the j slot: minus the force on i, summed over i slots
        double rl[3], rs[3];
        for (int q = 0; q < 3; ++q) {
            const double v = -force * d[q];
            rs[q] = isshort ? v : 0.0;
            rl[q] = isshort ? 0.0 : v;
        }
        for (int off = 8; off < 64; off <<= 1)
            for (int q = 0; q < 3; ++q) {
                rl[q] += __shfl_xor(rl[q], off);
                if (front) rs[q] += __shfl_xor(rs[q], off);
            }
        if (il == 0) {
            unsigned long long *w = a.jf + (size_t)sj * 8;
            for (int q = 0; q < 3; ++q) {
                const long long vl = __double2ll_rn(rl[q] * a.fscale);
                if (vl != 0) atomicAdd(w + 3 + q, (unsigned long long)vl);
                if (front) {
                    const long long vs = __double2ll_rn(rs[q] * a.fscale);
                    if (vs != 0) atomicAdd(w + q, (unsigned long long)vs);
                }
            }
        }
    }
    for (int off = 1; off < 8; off <<= 1)
        for (int q = 0; q < 3; ++q) {
            Fs[q] += __shfl_xor(Fs[q], off);
            Fl[q] += __shfl_xor(Fl[q], off);
        }
    if (jl == 0 && ai >= 0) {
        if (a.fs_x) { a.fs_x[ai] = Fs[0]; a.fs_y[ai] = Fs[1]; a.fs_z[ai] = Fs[2]; }
        if (a.fl_x) { a.fl_x[ai] = Fl[0]; a.fl_y[ai] = Fl[1]; a.fl_z[ai] = Fl[2]; }
    }
    double q[14] = {E[0], E[1], E[2], E[3], E[4], E[5], Vs[0], Vs[1], Vs[2], Vl[0], Vl[1], Vl[2],
                    (double)cnt_s, (double)cnt_l};
    for (int off = 1; off < 64; off <<= 1)
        for (int k = 0; k < 14; ++k) q[k] += __shfl_xor(q[k], off);
    if (lane == 0) {
        double *o = a.epart + (size_t)ci * Q_N;
        for (int k = 0; k < Q_N; ++k) o[k] = 0.0;
        o[Q_PELS2] = q[0];
        o[Q_PELS1] = q[1];
        o[Q_PE2] = q[2];
        o[Q_PE1] = q[3];
        o[Q_VIR2] = q[4];
        o[Q_VIR1] = q[5];
        o[Q_VIRFX] = q[6]; o[Q_VIRFY] = q[7]; o[Q_VIRFZ] = q[8];
        o[Q_VIRX] = q[9]; o[Q_VIRY] = q[10]; o[Q_VIRZ] = q[11];
        o[Q_NPS] = q[12];
        o[Q_NPL] = q[13];
    }
    if (DETAIL) {
        __syncthreads();
        detail_out(a, ci, s_det, s_ti);
    }
}

template <bool SOFT, bool DETAIL>
__global__ __launch_bounds__(64) void k_pairs64(const PairArgs *__restrict__ args, const int2 *__restrict__ map,
                                                int group) {
    extern __shared__ double s_det64[];   // detail buckets
    const int2 m = map[blockIdx.x];
    pairs64_body<SOFT, DETAIL>(args[m.x], m.y, s_det64, group);
}

template __global__ void k_pairs64<false, false>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs64<true, false>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs64<false, true>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs64<true, true>(const PairArgs *, const int2 *, int);

// j reactions per atom: one wavefront per j-cluster.  Every lane loads
// 16 B (four floats: four j slots of one component) of a record per step:
// the front-segment records (two units: long x, z, y, short x, y, z, 8 j
// slots each; lane = 12 (record of the step) + quarter, five records per
// step), then the long-class records (one unit: long x, z, y; lane = 6
// (record) + quarter, ten per step).  Each lane sums its floats as 64-bit
// fixed point (round(v fscale) as the low bits of v fscale + 1.5 2^52,
// exact for |v fscale| < 2^51, else converted directly), so the integer
// totals do not depend on the record order; the lanes' sums meet in LDS
// and lanes 0-47 each add up one (component, j slot) word, stored once:
// no atomics, no clearing.  A fast-class call (group 1) sums the short
// floats of the front-segment records only, a slow-class call (group 2)
// the long ones.
__device__ inline long long to_fixed(float v, double fscale, float vlim) {
    long long qv = __double_as_longlong(fma((double)v, fscale, kMagic)) - __double_as_longlong(kMagic);
    const bool big = fabsf(v) >= vlim;
    if (__ballot(big) != 0ull) {
        const long long slow = __double2ll_rn((double)v * fscale);
        qv = big ? slow : qv;
        asm volatile("" : "+v"(qv));    // a real branch: this path is rare
    }
    return qv;
}

__device__ inline void jgather_body(const PairArgs &a, int c, int group, long long *sh) {
    if (c >= a.nclu) return;
    const int lane = threadIdx.x;
    // front records: 12 quarters each, 5 per step (lanes 0-59)
    const int fq = lane % 12, fr = lane / 12;
    const bool fshort = fq >= 6;
    const bool fwant = fr < 5 && (fshort ? group != 2 : group != 1);
    // long records: 6 quarters each, 10 per step (lanes 0-59)
    const int lq = lane % 6, lr = lane / 6;
    const bool lwant = lr < 10 && group != 1;
    const unsigned u0 = sld(a.jstart)[c];
    const int2 nc = ldv(sld(a.jcnt) + c);
    using f4 = float __attribute__((ext_vector_type(4)));
    const MDX_GLOBAL f4 *rec = reinterpret_cast<const MDX_GLOBAL f4 *>(gld(a.rec) + (size_t)u0 * kRecUnit);
    const double fscale = a.fscale;
    const float vlim = (float)(2251799813685248.0 / fscale);   // 2^51 / fscale
    long long af[4] = {0, 0, 0, 0}, al[4] = {0, 0, 0, 0};
    auto add4 = [&](long long (&acc)[4], f4 v) {
        acc[0] += to_fixed(v.x, fscale, vlim);
        acc[1] += to_fixed(v.y, fscale, vlim);
        acc[2] += to_fixed(v.z, fscale, vlim);
        acc[3] += to_fixed(v.w, fscale, vlim);
    };
    constexpr int U = 4;       // 16-byte loads in flight per lane
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    for (int t0 = 0; t0 < nc.x; t0 += 5 * U) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = t0 + 5 * u + fr;
            v[u] = (fwant && r < nc.x) ? __builtin_nontemporal_load(rec + (size_t)r * 12 + fq) : z4;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) add4(af, v[u]);
    }
    if (group != 1) {
        const MDX_GLOBAL f4 *rl = rec + (size_t)nc.x * 12;
        for (int t0 = 0; t0 < nc.y; t0 += 10 * U) {
            f4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r = t0 + 10 * u + lr;
                v[u] = (lwant && r < nc.y) ? __builtin_nontemporal_load(rl + (size_t)r * 6 + lq) : z4;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) add4(al, v[u]);
        }
    }
    // sh[0..255]: front sums [lane][4], sh[256..511]: long sums [lane][4]
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sh[lane * 4 + k] = af[k];
        sh[256 + lane * 4 + k] = al[k];
    }
    __syncthreads();
    if (lane < 48) {
        // word: component comp (record rows: long x, z, y, short x, y, z)
        // and j slot jl; its floats are quarter 2 comp + jl / 4, element jl % 4
        const int comp = lane >> 3, jl = lane & 7;
        const int q = 2 * comp + (jl >> 2), e = jl & 3;
        long long t = 0;
        for (int r = 0; r < 5; ++r) t += sh[(12 * r + q) * 4 + e];
        if (comp < 3)
            for (int r = 0; r < 10; ++r) t += sh[256 + (6 * r + q) * 4 + e];
        constexpr int word[6] = {3, 5, 4, 0, 1, 2};   // long x, z, y -> 3, 5, 4; short x, y, z -> 0, 1, 2
        const bool want = comp < 3 ? group != 1 : group != 2;
        if (want) gst(a.jf)[((size_t)c * kClu + jl) * 8 + word[comp]] = (unsigned long long)t;
    }
}

__global__ __launch_bounds__(64) void k_jgather(const PairArgs *__restrict__ args, const int2 *__restrict__ map,
                                                int group) {
    __shared__ long long sh[512];
    const int2 m = map[blockIdx.x];
    jgather_body(args[m.x], m.y, group, sh);
}

// Per-slot LJ type, ee flag and the low part of the charge (at list build).
__global__ void k_slot_info(const int *__restrict__ pnclu, const int *__restrict__ clu_atom, const int *__restrict__ atom_ljtype,
                            const int *__restrict__ atom_moltype, const int *__restrict__ type_ee,
                            const double *__restrict__ q, int2 *__restrict__ aux) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= *pnclu * kClu) return;
    const int a = clu_atom[s];
    int2 v = make_int2(0, 0);
    if (a >= 0) {
        const float hi = (float)q[a];
        v.x = atom_ljtype[a] | (type_ee[atom_moltype[a]] << 16);
        v.y = __float_as_int((float)(q[a] - (double)hi));
    }
    aux[s] = v;
}

// One special pair per thread, reference expressions in FP64: LOCALF
// (forces.f:835-986) for the short classes, FORCES (1039-1198) for long.
template <bool DETAIL>
__device__ inline void special_body(const SpecialArgs &a, int blk, double *red, int group) {
    const int k = blk * blockDim.x + threadIdx.x;
    double q[Q_N];
    for (int i = 0; i < Q_N; ++i) q[i] = 0.0;
    int dt = -1;                 // molecule type of an evaluated pair (detail)
    double dcoul = 0.0, dlj = 0.0;
    if (k < a.ns) {
        const int c = a.cls[k];
        const bool use = c != SP_NONE && (group == 0 || (group == 1) == (c != SP_LONG));
        if (use) {
            const int isp = a.pairs[k].x, jsp = a.pairs[k].y;   // isp > jsp
            const int ityp = a.atom_moltype[isp];
            const int isn = a.atom_ljtype[isp], jsn = a.atom_ljtype[jsp];
            double qij, a6, b12;
            if (c == SP_SHORT14) {
                qij = a.q0[isp] * a.q0[jsp] * a.coulf * a.c14el[ityp];
                a6 = a.a6_14[isn * a.nlj + jsn] * a.c14lj[ityp];
                b12 = a.b12_14[isn * a.nlj + jsn] * a.c14lj[ityp];
            } else {
                qij = a.q0[isp] * a.q0[jsp] * a.coulf;
                a6 = a.a6_0[isn * a.nlj + jsn];
                b12 = a.b12_0[isn * a.nlj + jsn];
            }
            double dx = a.x[isp] - a.x[jsp], dy = a.y[isp] - a.y[jsp], dz = a.z[isp] - a.z[jsp];
            min_image(dx, dy, dz, a.box);
            double rr = dx * dx + dy * dy;
            rr = rr + dz * dz;
            if (c == SP_SHORT14 || rr <= a.rc2) {
                const double r1 = sqrt(rr);
                const bool localf = c != SP_LONG;
                double ee1, force;
                // special pairs are 1-4 pairs or pairs inside a lambda-coupled
                // molecule: direct Coulomb (forces.f:917-924, 1127-1134)
                if (localf) {
                    const double r1i = 1.0 / r1;
                    ee1 = qij * r1i;
                    force = ee1 * (r1i * r1i);
                } else {
                    ee1 = qij / r1;
                    force = ee1 / (r1 * r1);
                }
                const double ees = ee1;
                const double r1i = 1.0 / r1;
                const double r2i = r1i * r1i;
                const double r6i = r2i * r2i * r2i;
                const double esr = (a6 + b12 * r6i) * r6i;
                const double fsr = (6.0 * a6 + 12.0 * b12 * r6i) * r6i * r1i;
                if (b12 != 0.0) force = force + fsr * r1i;
                const double fxx = force * dx, fyy = force * dy, fzz = force * dz;
                if (DETAIL) {
                    // special pairs lie inside one molecule: PES141 / PSR141
                    dt = ityp;
                    dcoul = ee1;
                    dlj = b12 != 0.0 ? esr : 0.0;
                }
                if (localf) {
                    q[Q_NPS] = 1.0;
                    q[Q_PELS2] = ees;
                    if (b12 != 0.0) { q[Q_PE2] = esr; q[Q_VIR2] = fsr * r1; }
                    q[Q_VIRFX] = force * (dx * dx); q[Q_VIRFY] = force * (dy * dy); q[Q_VIRFZ] = force * (dz * dz);
                    fx_add(a.fixh_x + isp, fxx, a.fscale); fx_add(a.fixh_y + isp, fyy, a.fscale);
                    fx_add(a.fixh_z + isp, fzz, a.fscale);
                    fx_add(a.fixh_x + jsp, -fxx, a.fscale); fx_add(a.fixh_y + jsp, -fyy, a.fscale);
                    fx_add(a.fixh_z + jsp, -fzz, a.fscale);
                } else {
                    q[Q_NPL] = 1.0;
                    q[Q_PELS1] = ees;
                    if (b12 != 0.0) { q[Q_PE1] = esr; q[Q_VIR1] = fsr * r1; }
                    q[Q_VIRX] = force * (dx * dx); q[Q_VIRY] = force * (dy * dy); q[Q_VIRZ] = force * (dz * dz);
                    fx_add(a.fixg_x + isp, fxx, a.fscale); fx_add(a.fixg_y + isp, fyy, a.fscale);
                    fx_add(a.fixg_z + isp, fzz, a.fscale);
                    fx_add(a.fixg_x + jsp, -fxx, a.fscale); fx_add(a.fixg_y + jsp, -fyy, a.fscale);
                    fx_add(a.fixg_z + jsp, -fzz, a.fscale);
                }
            }
        }
    }
    // block sums in a fixed tree order
    for (int i = 0; i < Q_N; ++i) {
        red[threadIdx.x] = q[i];
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.epart[(size_t)blk * Q_N + i] = red[0];
        __syncthreads();
    }
    if (DETAIL) {
        double *o = a.dpart + (size_t)blk * a.ndet;
        for (int c = threadIdx.x; c < 2 * a.ntp; c += blockDim.x) o[c] = 0.0;
        for (int c = 0; c < 2 * a.nmt; ++c) {
            const int t = c < a.nmt ? c : c - a.nmt;
            red[threadIdx.x] = dt == t ? (c < a.nmt ? dcoul : dlj) : 0.0;
            __syncthreads();
            for (int s = blockDim.x / 2; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) o[2 * a.ntp + c] = red[0];
            __syncthreads();
        }
    }
}

template <bool DETAIL>
__global__ __launch_bounds__(256) void k_special(const SpecialArgs *__restrict__ args, const int2 *__restrict__ map,
                                                 int group) {
    __shared__ double red[256];
    const int2 m = map[blockIdx.x];
    special_body<DETAIL>(args[m.x], m.y, red, group);
}

template __global__ void k_special<false>(const SpecialArgs *, const int2 *, int);
template __global__ void k_special<true>(const SpecialArgs *, const int2 *, int);

// Column sums of the detail rows of one system: one block per column;
// thread t sums rows t, t + 256, ... in order, then a fixed tree over the
// 256 partial sums.
__global__ __launch_bounds__(256) void k_detail_reduce(const DetailArgs *__restrict__ args,
                                                       const int2 *__restrict__ map) {
    __shared__ double red[256];
    const int2 m = map[blockIdx.x];
    const DetailArgs a = args[m.x];
    const int c = m.y;
    double s = 0.0;
    for (int r = threadIdx.x; r < a.nrows; r += 256) s += a.dpart[(size_t)r * a.ndet + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.out[c] = red[0];
}

}  // namespace mdx
