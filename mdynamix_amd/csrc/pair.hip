// Real-space pair forces (replaces LOCALF forces.f:813-998 and FORCES
// forces.f:1014-1202).
//
// k_pairs: one wavefront per i-cluster of 8 atoms; each iteration takes one
// j-cluster of its list and lane (i_slot * 8 + j_slot) evaluates one atom
// pair in FP32: Lennard-Jones plus the Abramowitz-Stegun erfc-screened
// Ewald real-space term (the reference's polynomial, prcm.h:27-29, with the
// exact Gaussian in the force, forces.f:1146-1153).  Distance vectors are
// formed from anchor-relative FP32 coordinates plus the cluster pair's
// image shift, so no FP32 absolute coordinate enters a difference.
// The list is a full list (no Newton's third law): every atom's force is
// summed by the one wavefront that owns it, in a fixed order, with FP32
// partial sums over 16 j-clusters flushed into FP64; energies and virials
// (halved for the double visit) go to per-wavefront FP64 partials.  Pairs
// keep the short/long class of the list build: short pairs feed the fast
// forces H, long pairs the slow forces G.
//
// k_special: the 1-4 and lambda-coupled intramolecular pairs in FP64 with
// the reference's exact expressions (direct Coulomb, 1-4 scaling).
#include "mdx_internal.h"

namespace mdx {

struct PairArgs {
    const PairEntry *list;
    const int *nlist;
    int list_cap;
    const float4 *xq;            // per slot: x, y, z relative to anchor; q (FP32)
    const int2 *aux;             // per slot: LJ type | ee << 16; q - (float)q as float bits
    const double *q;             // per atom charge (FP64)
    float coulf_f;               // coulomb factor (FP32, forces)
    double coulf;                // coulomb factor (FP64, energies)
    const float2 *lj;            // nlj x nlj: (A6, B12)
    int nlj;
    const int *clu_atom;
    float alpha;
    float rc2;                   // evaluation cutoff^2 (Verlet) or +inf (reference)
    float raex;
    int wrap;                    // small box: re-apply the minimum image per pair
    float lx, ly, lz, ilx, ily, ilz;
    unsigned long long group_mask;   // restricts to short (FAST) or long (SLOW)
    int group;                   // 0 ALL, 1 FAST (short only), 2 SLOW (long only)
    double *fs_x, *fs_y, *fs_z;  // short-class pair forces (per atom)
    double *fl_x, *fl_y, *fl_z;  // long-class pair forces (per atom)
    double *epart;               // [nclu][Q_N]
};

template <bool SOFT>
__global__ __launch_bounds__(64) void k_pairs(PairArgs a) {
    __shared__ float2 s_lj[32 * 32];
    const int ci = blockIdx.x;
    const int lane = threadIdx.x;
    const int il = lane >> 3;
    const int nlj2 = a.nlj * a.nlj;
    for (int t = lane; t < nlj2; t += 64) s_lj[t] = a.lj[t];
    __syncthreads();

    const float4 pi = a.xq[ci * kClu + il];
    const int infoi = a.aux[ci * kClu + il].x;
    const int ti = infoi & 0xffff;
    const float qic = pi.w * a.coulf_f;
    const float2 *ljrow = s_lj + ti * a.nlj;
    const int jl = lane & 7;
    const float A1 = (float)kA1, A2 = (float)kA2, A3 = (float)kA3, A4 = (float)kA4, A5 = (float)kA5;
    const float B1 = (float)kB1, TOT = (float)kTOTPI;
    const float alpha = a.alpha;
    const float nal2 = -alpha * alpha * 1.44269504088896341f;   // -alpha^2 log2(e)
    const float talpha = TOT * alpha;

    double Fs[3] = {0, 0, 0}, Fl[3] = {0, 0, 0};
    double E[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // el_s, el_l, lj_s, lj_l, vlj_s, vlj_l, (unused)
    double Vs[3] = {0, 0, 0}, Vl[3] = {0, 0, 0};
    int cnt_s = 0, cnt_l = 0;
    const int n = a.nlist[ci];
    const PairEntry *L = a.list + (size_t)ci * a.list_cap;
    for (int k0 = 0; k0 < n; k0 += 16) {
        float fs[3] = {0.f, 0.f, 0.f}, fl[3] = {0.f, 0.f, 0.f};
        float es = 0.f, el = 0.f, ljs = 0.f, ljl = 0.f, vs = 0.f, vl = 0.f;
        float vxs = 0.f, vys = 0.f, vzs = 0.f, vxl = 0.f, vyl = 0.f, vzl = 0.f;
        const int k1 = min(k0 + 16, n);
        for (int k = k0; k < k1; ++k) {
            const PairEntry e = L[k];
            unsigned long long m = e.mask;
            if (a.group == 1) m &= e.mshort;
            else if (a.group == 2) m &= ~e.mshort;
            if (m == 0ull) continue;
            const float4 pj = a.xq[e.cj * kClu + jl];
            const int2 auxj = a.aux[e.cj * kClu + jl];
            const int infoj = auxj.x;
            const bool on = (m >> lane) & 1ull;
            const bool isshort = (e.mshort >> lane) & 1ull;
            float dx = (pi.x - pj.x) + e.sx;
            float dy = (pi.y - pj.y) + e.sy;
            float dz = (pi.z - pj.z) + e.sz;
            if (a.wrap) {
                dx = fmaf(-a.lx, rintf(dx * a.ilx), dx);
                dy = fmaf(-a.ly, rintf(dy * a.ily), dy);
                dz = fmaf(-a.lz, rintf(dz * a.ilz), dz);
            }
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!on || !(r2 <= a.rc2)) continue;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            float rei = rinv;
            if (SOFT) {
                const int eej = infoj >> 16, eei = infoi >> 16;
                if ((eei & eej) == 0) rei = __builtin_amdgcn_rcpf(r + a.raex);
            }
            const float ar = alpha * r;
            const float t = __builtin_amdgcn_rcpf(1.0f + B1 * ar);
            const float g = __builtin_amdgcn_exp2f(nal2 * r2);
            const float erfct = ((((A5 * t + A4) * t + A3) * t + A2) * t + A1) * t * g;
            const float ee1 = qic * pj.w * rei;
            // energy per unit q_i C with the exact q_j = hi + lo (the i side
            // is applied in FP64 at the end): no systematic charge rounding
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
            if (isshort) {
                fs[0] += fxx; fs[1] += fyy; fs[2] += fzz;
                es += ees; ljs += esr; vs += vr;
                vxs += fxx * dx; vys += fyy * dy; vzs += fzz * dz;
                ++cnt_s;
            } else {
                fl[0] += fxx; fl[1] += fyy; fl[2] += fzz;
                el += ees; ljl += esr; vl += vr;
                vxl += fxx * dx; vyl += fyy * dy; vzl += fzz * dz;
                ++cnt_l;
            }
        }
        for (int q = 0; q < 3; ++q) { Fs[q] += fs[q]; Fl[q] += fl[q]; }
        E[0] += es; E[1] += el; E[2] += ljs; E[3] += ljl; E[4] += vs; E[5] += vl;
        Vs[0] += vxs; Vs[1] += vys; Vs[2] += vzs;
        Vl[0] += vxl; Vl[1] += vyl; Vl[2] += vzl;
    }
    // i forces: sum over the 8 lanes of one i slot
    for (int off = 1; off < 8; off <<= 1)
        for (int q = 0; q < 3; ++q) {
            Fs[q] += __shfl_xor(Fs[q], off);
            Fl[q] += __shfl_xor(Fl[q], off);
        }
    const int ai = a.clu_atom[ci * kClu + il];
    const double qc = ai >= 0 ? a.q[ai] * a.coulf : 0.0;
    E[0] *= qc;
    E[1] *= qc;
    if (jl == 0 && ai >= 0) {
        if (a.fs_x) { a.fs_x[ai] = Fs[0]; a.fs_y[ai] = Fs[1]; a.fs_z[ai] = Fs[2]; }
        if (a.fl_x) { a.fl_x[ai] = Fl[0]; a.fl_y[ai] = Fl[1]; a.fl_z[ai] = Fl[2]; }
    }
    // energies and virials: wavefront sums, halved for the full list
    double q[14] = {E[0], E[1], E[2], E[3], E[4], E[5], Vs[0], Vs[1], Vs[2], Vl[0], Vl[1], Vl[2],
                    (double)cnt_s, (double)cnt_l};
    for (int off = 1; off < 64; off <<= 1)
        for (int k = 0; k < 14; ++k) q[k] += __shfl_xor(q[k], off);
    if (lane == 0) {
        double *o = a.epart + (size_t)ci * Q_N;
        for (int k = 0; k < Q_N; ++k) o[k] = 0.0;
        o[Q_PELS2] = 0.5 * q[0];
        o[Q_PELS1] = 0.5 * q[1];
        o[Q_PE2] = 0.5 * q[2];
        o[Q_PE1] = 0.5 * q[3];
        o[Q_VIR2] = 0.5 * q[4];
        o[Q_VIR1] = 0.5 * q[5];
        o[Q_VIRFX] = 0.5 * q[6]; o[Q_VIRFY] = 0.5 * q[7]; o[Q_VIRFZ] = 0.5 * q[8];
        o[Q_VIRX] = 0.5 * q[9]; o[Q_VIRY] = 0.5 * q[10]; o[Q_VIRZ] = 0.5 * q[11];
        o[Q_NPS] = 0.5 * q[12];
        o[Q_NPL] = 0.5 * q[13];
    }
}

template __global__ void k_pairs<false>(PairArgs);
template __global__ void k_pairs<true>(PairArgs);

// Anchor-relative packed coordinates of every cluster slot (each step).
__global__ void k_pack(int nslot, const int *__restrict__ clu_atom, const double *__restrict__ anchor,
                       const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                       const double *__restrict__ q, Box b, float4 *__restrict__ xq) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int a = clu_atom[s];
    const int c = s / kClu;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a >= 0) {
        double d[3] = {x[a] - anchor[c * 3], y[a] - anchor[c * 3 + 1], z[a] - anchor[c * 3 + 2]};
        for (int k = 0; k < 3; ++k) d[k] -= b.l[k] * rint(d[k] / b.l[k]);
        o = make_float4((float)d[0], (float)d[1], (float)d[2], (float)q[a]);
    }
    xq[s] = o;
}

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

struct SpecialArgs {
    int ns;
    const int2 *pairs;
    const signed char *cls;
    const double *x, *y, *z;
    const int *atom_moltype, *atom_ljtype;
    const double *q0;            // CHARGE(NSITE)
    const int *type_ee;
    const double *c14el, *c14lj;
    const double *a6_0, *b12_0, *a6_14, *b12_14;
    int nlj;
    double coulf, alpha, rc2;    // rc2: evaluation cutoff (Verlet) or +inf
    Box box;
    double fscale;
    unsigned long long *fixh_x, *fixh_y, *fixh_z;   // short / 1-4 -> H (LOCALF)
    unsigned long long *fixg_x, *fixg_y, *fixg_z;   // long -> G (FORCES)
    int group;
    double *epart;               // [gridDim][Q_N]
};

// One special pair per thread, reference expressions in FP64: LOCALF
// (forces.f:835-986) for the short classes, FORCES (1039-1198) for long.
__global__ __launch_bounds__(256) void k_special(SpecialArgs a) {
    __shared__ double red[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double q[Q_N];
    for (int i = 0; i < Q_N; ++i) q[i] = 0.0;
    if (k < a.ns) {
        const int c = a.cls[k];
        const bool use = c != SP_NONE && (a.group == 0 || (a.group == 1) == (c != SP_LONG));
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
                double ee1, ees, force;
                const bool direct = c == SP_SHORT14 || a.type_ee[ityp] == 0;
                if (direct) {
                    if (localf) {
                        const double r1i = 1.0 / r1;
                        ee1 = qij * r1i;
                        force = ee1 * (r1i * r1i);
                    } else {
                        ee1 = qij / r1;
                        force = ee1 / (r1 * r1);
                    }
                    ees = ee1;
                } else {
                    const double r1i = 1.0 / r1;
                    const double alphar = a.alpha * r1;
                    const double ttt = 1.0 / (1.0 + kB1 * alphar);
                    const double exp2a = exp(-(alphar * alphar));
                    const double erfct = ((((kA5 * ttt + kA4) * ttt + kA3) * ttt + kA2) * ttt + kA1) * ttt * exp2a;
                    ee1 = qij * r1i;
                    ees = ee1 * erfct;
                    force = ee1 * (kTOTPI * a.alpha * exp2a + erfct * r1i) * r1i;
                }
                const double r1i = 1.0 / r1;
                const double r2i = r1i * r1i;
                const double r6i = r2i * r2i * r2i;
                const double esr = (a6 + b12 * r6i) * r6i;
                const double fsr = (6.0 * a6 + 12.0 * b12 * r6i) * r6i * r1i;
                if (b12 != 0.0) force = force + fsr * r1i;
                const double fxx = force * dx, fyy = force * dy, fzz = force * dz;
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
        if (threadIdx.x == 0) a.epart[(size_t)blockIdx.x * Q_N + i] = red[0];
        __syncthreads();
    }
}

}  // namespace mdx
