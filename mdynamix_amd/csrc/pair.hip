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
// x(lane ^ 8): a rotation by 8 inside each row of 16 (DPP row_ror:8).
__device__ inline float xor8(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x128, 0xf, 0xf, false));
}

// Detail calls (MDX_WANT_DETAIL) also split the bare Coulomb energy
// q_i q_j C / (r + RAE) and the LJ energy by molecule-type pair (POTE1,
// POTL1) or, for pairs inside one molecule, by molecule type (PES141,
// PSR141), as LOCALF/FORCES do (forces.f:945-974, 1155-1183).  Each lane
// keeps its own FP64 buckets in LDS (bucket = type of the j molecule, or
// the intramolecular bucket), so the sums stay in a fixed order.
template <bool SOFT, bool DETAIL>
__device__ inline void pairs_body(const PairArgs &a, int ci, float2 *s_lj, float *s_jf, int *s_jc, double *s_acc,
                                  int group) {
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
    const float4 pi = ldv(gld(a.xq) + ci * kClu + il);
    const int infoi = gld(a.aux)[ci * kClu + il].x;
    const int ti = infoi & 0xffff;
    const float2 *ljrow = s_lj + ti * a.nlj;
    const int jl = lane & 7;
    const float A1 = (float)kA1, A2 = (float)kA2, A3 = (float)kA3, A4 = (float)kA4, A5 = (float)kA5;
    const float B1 = (float)kB1, TOT = (float)kTOTPI;
    const float alpha = a.alpha;
    const float nal2 = -alpha * alpha * 1.44269504088896341f;   // -alpha^2 log2(e)
    const float talpha = TOT * alpha;
    const float qic = pi.w * a.coulf_f;
    int moli = -1;
    if (DETAIL) {
        const int ai0 = a.clu_atom[ci * kClu + il];
        moli = ai0 >= 0 ? a.atom_mol[ai0] : -1;
        if (jl == 0) s_ti[il] = ai0 >= 0 ? a.atom_moltype[ai0] : 0;
    }

    double Fs[3] = {0, 0, 0}, Fl[3] = {0, 0, 0};
    double E[6] = {0, 0, 0, 0, 0, 0};   // el_s, el_l, lj_s, lj_l, vlj_s, vlj_l
    // the FP64 virial sums per lane live in LDS (touched once per block of
    // 16 entries; registers stay free for the pair loop): short xyz, long xyz
    for (int q = 0; q < 6; ++q) s_acc[q * 64 + lane] = 0.0;
    int cnt_s = 0, cnt_l = 0;   // wave totals (ballot counts)
    // a fast-class call reads only the entries with a short-class pair (the
    // front segment); logical entry k >= ns sits at list_cap - 1 - (k - ns)
    const int ns = sld(a.nshort)[ci];
    const int n = group == 1 ? ns : sld(a.nlist)[ci];
    const int cap = a.list_cap;
    auto ent = [ns, cap](int k) { return k < ns ? k : cap - 1 - (k - ns); };
    const MDX_CONST PairEntry *L = sld(a.list) + (size_t)ci * a.list_cap;
    const MDX_GLOBAL float4 *xq = gld(a.xq);
    const MDX_GLOBAL int2 *aux = gld(a.aux);
    MDX_GLOBAL unsigned long long *const jf_base = gst(a.jf) + (size_t)jl * 8 + il;
    const double fscale = a.fscale;
    const float bl[3] = {a.lx, a.ly, a.lz}, bi[3] = {a.ilx, a.ily, a.ilz};
    const float rc2 = a.rc2, raex = a.raex, band = a.band;
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
    // software pipeline: list entries two ahead (scalar loads), the
    // j-cluster coordinates one ahead
    PairEntry e1{}, e2{};
    float4 pj1 = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 aux1 = make_int2(0, 0);
    if (n > 0) e1 = ldv(L + ent(0));
    if (n > 1) e2 = ldv(L + ent(1));
    auto jclu = [](const PairEntry &e) { return (int)((unsigned)e.cj & ~kEntryWrap); };
    if (n > 0) { pj1 = ldv(xq + jclu(e1) * kClu + jl); aux1 = ldv(aux + jclu(e1) * kClu + jl); }
    for (int k0 = 0; k0 < n; k0 += 16) {
        float fs[3] = {0.f, 0.f, 0.f}, fl[3] = {0.f, 0.f, 0.f};
        float es = 0.f, el = 0.f, ljs = 0.f, ljl = 0.f, vs = 0.f, vl = 0.f;
        float vxs = 0.f, vys = 0.f, vzs = 0.f, vxl = 0.f, vyl = 0.f, vzl = 0.f;
        const int k1 = min(k0 + 16, n);
        // one list entry; MIXED: an entry of the front segment (it holds a
        // short-class pair, so both classes are split by exact 0/1 weights);
        // otherwise every pair is long-class and only the long sums move
        auto entry = [&](auto mixed, int kk, const PairEntry &e, const float4 &pj, const int2 &auxj) {
            constexpr bool MIXED = decltype(mixed)::value;
            unsigned long long m = e.mask;
            if (group == 1) m &= e.mshort;
            else if (group == 2) m &= ~e.mshort;
            // branch-free body: a lane without a pair inside the cutoff
            // takes rinv = 0, which makes every contribution exactly zero
            // (no divergent paths to merge)
            const int infoj = auxj.x;
            const bool on = (m >> lane) & 1ull;
            const bool isshort = MIXED && ((e.mshort >> lane) & 1ull);
            const float ws = isshort ? 1.f : 0.f, wl = 1.f - ws;
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
            const float r2 = dx * dx + dy * dy + dz * dz;
            bool in = on && r2 <= rc2;
            if (band > 0.f) {
                // near the cutoff: the reference's FP64 test decides
                const bool amb = on && fabsf(r2 - rc2) <= band;
                if (__ballot(amb) != 0ull && amb) in = exact_in(jclu(e));
            }
            const float rinv = in ? __builtin_amdgcn_rsqf(r2) : 0.f;
            const float r = r2 * rinv;
            float rei = rinv;
            if (SOFT) {
                const int eej = infoj >> 16, eei = infoi >> 16;
                rei = (in && (eei & eej) == 0) ? __builtin_amdgcn_rcpf(r + raex) : rinv;
            }
            const float ar = alpha * r;
            const float t = __builtin_amdgcn_rcpf(1.0f + B1 * ar);
            const float g = __builtin_amdgcn_exp2f(nal2 * r2);
            const float erfct = ((((A5 * t + A4) * t + A3) * t + A2) * t + A1) * t * g;
            const float ee1 = qic * pj.w * rei;
            const float phi = erfct * rei;
            const float ees = pj.w * phi + __int_as_float(auxj.y) * phi;
            const float2 lj = ljrow[infoj & 0xffff];
            const float r2i = rei * rei;
            const float r6i = r2i * r2i * r2i;
            const float esr = (lj.x + lj.y * r6i) * r6i;
            const float fsr = (6.0f * lj.x + 12.0f * lj.y * r6i) * r6i * rei;
            const float fsc = (ee1 * (talpha * g + erfct * rei) + fsr) * rinv;
            const float fxx = fsc * dx, fyy = fsc * dy, fzz = fsc * dz;
            const float vr = fsr * r;
            if (DETAIL && in) {
                const int aj = a.clu_atom[jclu(e) * kClu + jl];
                const int b = a.atom_mol[aj] == moli ? a.nmt : a.atom_moltype[aj];
                s_det[(b * 2) * 64 + lane] += (double)(pj.w * rei) + (double)(__int_as_float(auxj.y) * rei);
                s_det[(b * 2 + 1) * 64 + lane] += (double)esr;
            }
            const float vxx = fxx * dx, vyy = fyy * dy, vzz = fzz * dz;
            const unsigned long long inm = __ballot(in);
            if (MIXED) {
                fs[0] = fmaf(ws, fxx, fs[0]); fs[1] = fmaf(ws, fyy, fs[1]); fs[2] = fmaf(ws, fzz, fs[2]);
                fl[0] = fmaf(wl, fxx, fl[0]); fl[1] = fmaf(wl, fyy, fl[1]); fl[2] = fmaf(wl, fzz, fl[2]);
                es = fmaf(ws, ees, es); ljs = fmaf(ws, esr, ljs); vs = fmaf(ws, vr, vs);
                el = fmaf(wl, ees, el); ljl = fmaf(wl, esr, ljl); vl = fmaf(wl, vr, vl);
                vxs = fmaf(ws, vxx, vxs); vys = fmaf(ws, vyy, vys); vzs = fmaf(ws, vzz, vzs);
                vxl = fmaf(wl, vxx, vxl); vyl = fmaf(wl, vyy, vyl); vzl = fmaf(wl, vzz, vzl);
                cnt_s += __popcll(inm & e.mshort);
                cnt_l += __popcll(inm & ~e.mshort);
            } else {
                fl[0] += fxx; fl[1] += fyy; fl[2] += fzz;
                el += ees; ljl += esr; vl += vr;
                vxl += vxx; vyl += vyy; vzl += vzz;
                cnt_l += __popcll(inm);
            }
            // Newton's third law: the j reactions (short x, y, z, long x, y,
            // z, two zeros) summed over the 8 i slots by a reduce-scatter
            // butterfly; lane (il, jl) ends with component il of j slot jl
            float w[4], u[2];
            if (MIXED) {
                w[0] = swap_sum32(-ws * fxx, -wl * fyy);
                w[1] = swap_sum32(-ws * fyy, -wl * fzz);
                w[2] = swap_sum32(-ws * fzz, 0.f);
                w[3] = swap_sum32(-wl * fxx, 0.f);
            } else {
                w[0] = swap_sum32(0.f, -fyy);
                w[1] = swap_sum32(0.f, -fzz);
                w[2] = 0.f;
                w[3] = swap_sum32(-fxx, 0.f);
            }
            for (int q = 0; q < 2; ++q) u[q] = swap_sum16(w[q], w[2 + q]);
            const float send = (lane & 8) ? u[0] : u[1], keep = (lane & 8) ? u[1] : u[0];
            s_jf[kk * 64 + lane] = keep + xor8(send);
            if (lane == 0) s_jc[kk] = jclu(e);
        };
        for (int k = k0; k < k1; ++k) {
            const PairEntry e = e1;
            const float4 pj = pj1;
            int2 auxj = aux1;
            // keeps the LJ-type mask below from being hoisted next to the
            // prefetch load, which would wait for that load at once
            asm volatile("" : "+v"(auxj.x));
            e1 = e2;
            // unconditional prefetch (past the end it re-reads the last
            // entry and its j-cluster: valid addresses, values unused)
            e2 = ldv(L + ent(min(k + 2, n - 1)));
            pj1 = ldv(xq + jclu(e1) * kClu + jl);
            aux1 = ldv(aux + jclu(e1) * kClu + jl);
            if (k < ns) entry(std::true_type{}, k - k0, e, pj, auxj);
            else entry(std::false_type{}, k - k0, e, pj, auxj);
        }
        // the block's j reactions into the 64-bit fixed-point accumulators:
        // one atomic per (entry, component, j slot), batched so the wave
        // waits for them once per block
        __syncthreads();
        for (int k = k0; k < k1; ++k) {
            const float r = s_jf[(k - k0) * 64 + lane];
            if (il < 6 && r != 0.f) fx_add_g(jf_base + (size_t)s_jc[k - k0] * (kClu * 8), (double)r, fscale);
        }
        __syncthreads();
        for (int q = 0; q < 3; ++q) { Fs[q] += fs[q]; Fl[q] += fl[q]; }
        E[0] += es; E[1] += el; E[2] += ljs; E[3] += ljl; E[4] += vs; E[5] += vl;
        {
            const float part[6] = {vxs, vys, vzs, vxl, vyl, vzl};
            for (int q = 0; q < 6; ++q) s_acc[q * 64 + lane] += (double)part[q];
        }
    }
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
        // bare Coulomb buckets per unit q_i C -> energies; then each output
        // column sums its (lane, bucket) entries in lane order; the pair
        // type index is symmetric, so each pair counts once in its bucket
        for (int b = 0; b <= a.nmt; ++b) s_det[(b * 2) * 64 + lane] *= qc;
        __syncthreads();
        for (int o = lane; o < a.ndet; o += 64) {
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
}

template <bool SOFT, bool DETAIL>
__global__ __launch_bounds__(64) void k_pairs(const PairArgs *__restrict__ args, const int2 *__restrict__ map,
                                              int group) {
    extern __shared__ float2 s_lj[];      // nlj x nlj (A6, B12), then detail buckets
    __shared__ float s_jf[16 * 64];       // j reactions of a block of 16 list entries
    __shared__ int s_jc[16];              // their j-clusters
    __shared__ double s_acc[6 * 64];      // per-lane FP64 virial sums
    const int2 m = map[blockIdx.x];
    pairs_body<SOFT, DETAIL>(args[m.x], m.y, s_lj, s_jf, s_jc, s_acc, group);
}

template __global__ void k_pairs<false, false>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<true, false>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<false, true>(const PairArgs *, const int2 *, int);
template __global__ void k_pairs<true, true>(const PairArgs *, const int2 *, int);

// Per-slot LJ type, ee flag and the low part of the charge (at list build).
__global__ void k_slot_info(int nslot, const int *__restrict__ clu_atom, const int *__restrict__ atom_ljtype,
                            const int *__restrict__ atom_moltype, const int *__restrict__ type_ee,
                            const double *__restrict__ q, int2 *__restrict__ aux) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
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

// Column sums of the detail rows of one system, in row order.
__global__ __launch_bounds__(256) void k_detail_reduce(const DetailArgs *__restrict__ args,
                                                       const int2 *__restrict__ map) {
    const DetailArgs a = args[map[blockIdx.x].x];
    for (int c = threadIdx.x; c < a.ndet; c += blockDim.x) {
        double s = 0.0;
        for (int r = 0; r < a.nrows; ++r) s += a.dpart[(size_t)r * a.ndet + c];
        a.out[c] = s;
    }
}

}  // namespace mdx
