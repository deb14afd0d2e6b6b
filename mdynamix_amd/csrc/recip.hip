// Ewald reciprocal sum (replaces FURIR, forces.f:1217-1339).
//
// The half-space k set of INIFUR (forces.f:1356-1427) is regrouped into
// rows of fixed (n_x, n_y); within a row the vectors n_z = +j and -j are
// taken together (the k set is symmetric in n_z except the (0, 0) row).
// k_recip_prep: per-atom phase tables along x, y and z (FP64 recurrences);
// k_sk: S(+-j) = U_j +- i V_j, with U_j, V_j sums of B p_j over the atoms
//   of a chunk (B = q exp(2 pi i (n_x f_x + n_y f_y)), p_j = exp(2 pi i j
//   f_z)), as a matrix product on the matrix cores whose result for each
//   32-atom stage is added into FP64 sums;
// k_sk_reduce: FP64 sums of the chunk partials, energy, virial, and the
//   per-pair data of the force pass;
// k_recip_force: lane = atom, rows uniform, the force
//   q sum_k k Im(exp(i k.r) conj(A_k)) (forces.f:1320-1326) evaluated for
//   +-j together (see recip_force_body).
#include "mdx_internal.h"

namespace mdx {

// f16 pair of an FP32 value: hi = f16(v), lo = f16(v - hi); hi + lo keeps
// 22 significant bits (absolute error below 2^-24 for |v| <= 1, the range
// of B and p_j), and the three products hi*hi + hi*lo + lo*hi of two such
// pairs reproduce an FP32 product to about 2^-21 relative.
//
// The value is taken through an empty asm first.  Otherwise FP contraction
// may fuse the product that formed v into both uses with different
// roundings (hi from the FP32-rounded product, lo = f16(exact product -
// hi) in one mixed FMA); where the two sit on either side of an f16
// rounding midpoint, hi + lo then misses v by a whole f16 ulp.
__device__ inline void split1(float v, _Float16 &hi, _Float16 &lo) {
    asm volatile("" : "+v"(v));
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

__device__ inline void recip_prep_body(const RecipArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    const bool have = i < a.natoms;
    const size_t n = (size_t)a.tstride;
    double c1 = 1.0, s1 = 0.0;
    if (!have) {
        // the tables' padding atoms are zero (the structure-factor pass
        // reads atoms in pairs and stages up to the padded row end)
        if (i < a.tstride) {
            for (int k = 0; k <= a.kx; ++k) stv(gst(a.tx) + i + k * n, make_float2(0.f, 0.f));
            for (int k = 0; k <= a.ky; ++k) stv(gst(a.ty) + i + k * n, make_float2(0.f, 0.f));
        }
    } else {
        const double fx = a.x[i] / a.box.l[0], fy = a.y[i] / a.box.l[1], fz = a.z[i] / a.box.l[2];
        const double q = a.q[i];
        stv(gst(a.fq) + i, make_double4(fx, fy, fz, q));
        // exp(2 pi i f_z) and its 8th power, in FP64
        const double u = fz - rint(fz);
        double s8, c8;
        sincospi(2.0 * u, &s1, &c1);
        double u8 = 8.0 * u;
        u8 -= rint(u8);
        sincospi(2.0 * u8, &s8, &c8);
        stv(gst(a.stepz) + i, make_double4(c1, s1, c8, s8));
        // phase tables (FP64 recurrences from FP64 sincospi steps, FP32 values)
        {
            double sr = 0.0, si = 0.0;
            sincospi(2.0 * (fx - rint(fx)), &si, &sr);
            double pr = q, pi = 0.0;
            MDX_GLOBAL float2 *t = gst(a.tx) + i;
            for (int k = 0; k <= a.kx; ++k) {
                stv(t + k * n, make_float2((float)pr, (float)pi));
                const double w = pr * sr - pi * si;
                pi = pr * si + pi * sr;
                pr = w;
            }
        }
        {
            double sr = 0.0, si = 0.0;
            sincospi(2.0 * (fy - rint(fy)), &si, &sr);
            double pr = 1.0, pi = 0.0;
            MDX_GLOBAL float2 *t = gst(a.ty) + i;
            for (int k = 0; k <= a.ky; ++k) {
                stv(t + k * n, make_float2((float)pr, (float)pi));
                const double w = pr * sr - pi * si;
                pi = pr * si + pi * sr;
                pr = w;
            }
        }
    }
    // z table, already split for the matrix instructions: per (j, atom
    // pair) 16 bytes = half2 (C hi), (D hi), (C lo), (D lo) of the pair's
    // two atoms, p_j = C + i D = exp(2 pi i j f_z) from the FP64 recurrence
    // rounded to FP32, hi = f16(v), lo = f16(v - hi); the even lane of each
    // pair writes the record (every lane of the wavefront takes part)
    {
        double pr = 1.0, pi = 0.0;
        uint4 *t = reinterpret_cast<uint4 *>(a.tz);
        const bool odd = i & 1;
        for (int k = 0; k < a.mzp; ++k) {
            const float c = have ? (float)pr : 0.f, d = have ? (float)pi : 0.f;
            _Float16 hc, hd, lc, ld;
            split1(c, hc, lc);
            split1(d, hd, ld);
            const unsigned hi2 = (unsigned)__builtin_bit_cast(unsigned short, hc) |
                                 ((unsigned)__builtin_bit_cast(unsigned short, hd) << 16);
            const unsigned lo2 = (unsigned)__builtin_bit_cast(unsigned short, lc) |
                                 ((unsigned)__builtin_bit_cast(unsigned short, ld) << 16);
            const unsigned ohi = __shfl_xor(hi2, 1), olo = __shfl_xor(lo2, 1);
            if (!odd && i < a.tstride) {
                uint4 r;
                r.x = (hi2 & 0xffffu) | (ohi << 16);            // C hi of atoms i, i + 1
                r.y = (hi2 >> 16) | (ohi & 0xffff0000u);         // D hi
                r.z = (lo2 & 0xffffu) | (olo << 16);            // C lo
                r.w = (lo2 >> 16) | (olo & 0xffff0000u);         // D lo
                stv(gst(t) + (k * n + i) / 2, r);
            }
            const double w = pr * c1 - pi * s1;
            pi = pr * s1 + pi * c1;
            pr = w;
        }
    }
}

__global__ __launch_bounds__(256) void k_recip_prep(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    recip_prep_body(args[m.x], m.y);
}

typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// f16 hi/lo pairs of two FP32 values (split1 each).
__device__ inline void split_f16(float v0, float v1, half2_t &hi, half2_t &lo) {
    _Float16 h0, h1, l0, l1;
    split1(v0, h0, l0);
    split1(v1, h1, l1);
    hi = half2_t{h0, h1};
    lo = half2_t{l0, l1};
}

// LDS stage of the structure-factor block: per plane row (a k row or a
// value of j, times Re/Im or C/D) 32 atoms of f16 (64 B) plus 16 B of
// padding; four planes per stage: W hi, W lo, Z hi, Z lo.
constexpr int kSkRowB = 80;                 // bytes per plane row

template <int NT>
struct SkStage {
    static constexpr int WROWS = kSkRows * 2;
    static constexpr int ZROWS = NT * 8 * 2;
    static constexpr int WH = 0, WL = WROWS * kSkRowB, ZH = 2 * WROWS * kSkRowB, ZL = ZH + ZROWS * kSkRowB;
    static constexpr int BYTES = ZL + ZROWS * kSkRowB;
};

// Structure factors on the matrix cores.  For a row (n_x, n_y) with
// B_a = q_a exp(2 pi i (n_x f_x + n_y f_y)) and p_j = exp(2 pi i j f_z) =
// C_j + i D_j, the vectors n_z = +-j have S(+-j) = U_j +- i V_j with
// U_j = sum_a B_a C_j(a), V_j = sum_a B_a D_j(a); n_z = 0 is U_0 (C_0 = 1).
// Over the atoms of a chunk this is the real product
//   [Re B; Im B] (rows x atoms)  x  [C_j, D_j] (atoms x j)
// evaluated with v_mfma_f32_16x16x32_f16 on f16 hi/lo pairs of the FP32
// operands (three products per tile, hi*hi + hi*lo + lo*hi): an M tile is
// 8 rows x (Re, Im), an N tile 8 values of j x (C, D), K = the 32 atoms of
// an LDS stage.  The instruction sums its 32 atoms' products in FP32; each
// stage's result is added into FP64 accumulators, so the sum over the
// chunk's atoms is an FP64 sum of 32-atom partials.  A block holds 64 rows (sorted by pair count; a wavefront
// takes 16 rows = 2 M tiles and as many N tiles as its first, longest row
// needs) and one chunk of atoms, staged through LDS 32 atoms at a time:
// B = tx[n_x] ty[|n_y|] (conjugated for n_y < 0) formed in FP32, then
// split; p_j = tz[j] comes split from k_recip_prep.  Operand layout of the instruction (lane
// l): A[l % 16][8 (l / 16) + k], B[8 (l / 16) + k][l % 16], k = 0..7;
// result D[4 (l / 16) + v][l % 16].  Stages alternate between two LDS
// buffers, the next one fetched into registers while the current one feeds
// the matrix instructions.  Each block writes its FP64 chunk sums per pair
// entry (U_j, V_j) and per row (S(n_z = 0)); k_sk_reduce adds the chunks in
// FP64.  Blocks are numbered so that the row groups of one chunk share an
// XCD (block index mod 8), whose L2 then serves that chunk's table columns.
template <int NT>
__device__ inline void sk_body(const RecipArgs &a, int blk, unsigned char *lds) {
    using St = SkStage<NT>;
    const int G = a.nsgroups;
    const int x8 = blk & 7, s8 = blk >> 3;
    const int g = s8 % G, ch = (s8 / G) * 8 + x8;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int a0 = ch * a.chunk;
    const int a1 = min(a0 + a.chunk, a.natoms);
    const int rbase = g * kSkRows;
    // staging assignment: atom pair tid % 16 (atoms 2p, 2p + 1 of the
    // stage), rows tid / 16 + 16 i
    const int sp = tid & 15, sr0 = tid >> 4;
    int offx[4], offy[4];
    bool conj[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // rows past the last one stage row 0's tables; their sums are
        // never written
        const int r = rbase + sr0 + 16 * i;
        offx[i] = 0;
        offy[i] = 0;
        conj[i] = false;
        if (r < a.nfrows) {
            const FRow fr = ldv(gld(a.srows) + r);
            offx[i] = fr.ix * a.tstride;
            offy[i] = abs(fr.iy) * a.tstride;
            conj[i] = fr.iy < 0;
        }
    }
    // N tiles: the group's first row is its longest, a wavefront's first row is its longest
    const int rw = rbase + 16 * w;
    const int ntw = rw < a.nfrows ? min(NT, (ldv(gld(a.srows) + rw).m + 8) >> 3) : 0;
    const int ntg = rbase < a.nfrows ? min(NT, (ldv(gld(a.srows) + rbase).m + 8) >> 3) : 0;
    double acc[2][NT][4];           // FP64 sums over the chunk
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[mt][t][v] = 0.0;
    const mfma_f4 zero4 = {0.f, 0.f, 0.f, 0.f};
    const MDX_GLOBAL float4 *tx = reinterpret_cast<const MDX_GLOBAL float4 *>(gld(a.tx));
    const MDX_GLOBAL float4 *ty = reinterpret_cast<const MDX_GLOBAL float4 *>(gld(a.ty));
    const MDX_GLOBAL float4 *tz = reinterpret_cast<const MDX_GLOBAL float4 *>(gld(a.tz));
    // Z items of a stage: (j, atom pair) for j < 8 ntg, tid + 256 k
    constexpr int ZIT = (NT * 8 * 16 + 255) / 256;
    float4 X[4], Y[4], Zv[ZIT];
    // Unconditional loads: a chunk's stages never pass the chunk's end
    // except in the last chunk, whose last stage ends at or before the
    // 64-atom padded row end, and the padding is zero; rows of j past the
    // group's N tiles re-read its last row (no matrix instruction uses them).
    const int jlast = max(0, 8 * ntg - 1);
    // per-thread byte offsets (16 B = two atoms) are fixed for the whole
    // chunk; a stage only moves the uniform base (scalar base + 32-bit
    // vector offset addressing; every table is far below 4 GiB)
    unsigned ox[4], oy[4], oz[ZIT];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ox[i] = ((unsigned)offx[i] / 2 + sp) * 16u;
        oy[i] = ((unsigned)offy[i] / 2 + sp) * 16u;
    }
#pragma unroll
    for (int k = 0; k < ZIT; ++k) {
        const int it = tid + 256 * k;
        oz[k] = ((unsigned)min(it >> 4, jlast) * ((unsigned)a.tstride / 2) + (it & 15)) * 16u;
    }
    auto at = [](const MDX_GLOBAL float4 *base, unsigned off) {
        return ldv(reinterpret_cast<const MDX_GLOBAL float4 *>(reinterpret_cast<const MDX_GLOBAL char *>(base) + off));
    };
    auto fetch = [&](int sb) {
        const MDX_GLOBAL float4 *txs = tx + sb / 2, *tys = ty + sb / 2, *tzs = tz + sb / 2;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            X[i] = at(txs, ox[i]);
            Y[i] = at(tys, oy[i]);
        }
#pragma unroll
        for (int k = 0; k < ZIT; ++k) Zv[k] = at(tzs, oz[k]);
    };
    auto store = [&](unsigned char *buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = sr0 + 16 * i;
            const float y1 = conj[i] ? -Y[i].y : Y[i].y, y3 = conj[i] ? -Y[i].w : Y[i].w;
            const float re0 = X[i].x * Y[i].x - X[i].y * y1, im0 = X[i].x * y1 + X[i].y * Y[i].x;
            const float re1 = X[i].z * Y[i].z - X[i].w * y3, im1 = X[i].z * y3 + X[i].w * Y[i].z;
            half2_t hr, lr, hi, li;
            split_f16(re0, re1, hr, lr);
            split_f16(im0, im1, hi, li);
            const int o = (2 * r) * kSkRowB + 4 * sp;
            *reinterpret_cast<half2_t *>(buf + St::WH + o) = hr;
            *reinterpret_cast<half2_t *>(buf + St::WH + o + kSkRowB) = hi;
            *reinterpret_cast<half2_t *>(buf + St::WL + o) = lr;
            *reinterpret_cast<half2_t *>(buf + St::WL + o + kSkRowB) = li;
        }
#pragma unroll
        for (int k = 0; k < ZIT; ++k) {
            const int it = tid + 256 * k;
            const int j = it >> 4, p = it & 15;
            if (j < NT * 8) {
                // the z table is stored split (k_recip_prep): C hi, D hi, C lo, D lo
                const half2_t hc = __builtin_bit_cast(half2_t, Zv[k].x), hd = __builtin_bit_cast(half2_t, Zv[k].y);
                const half2_t lc = __builtin_bit_cast(half2_t, Zv[k].z), ld = __builtin_bit_cast(half2_t, Zv[k].w);
                const int o = (2 * j) * kSkRowB + 4 * p;
                *reinterpret_cast<half2_t *>(buf + St::ZH + o) = hc;
                *reinterpret_cast<half2_t *>(buf + St::ZH + o + kSkRowB) = hd;
                *reinterpret_cast<half2_t *>(buf + St::ZL + o) = lc;
                *reinterpret_cast<half2_t *>(buf + St::ZL + o + kSkRowB) = ld;
            }
        }
    };
    const int rA = lane & 7, cA = (lane >> 3) & 1, kA = lane >> 4;
    const int nsub = a1 > a0 ? (a1 - a0 + kSkSub - 1) / kSkSub : 0;
    if (nsub > 0) {
        fetch(a0);
        store(lds);
    }
    __syncthreads();
    for (int k = 0; k < nsub; ++k) {
        const unsigned char *cur = lds + (k & 1) * St::BYTES;
        // the last iteration re-fetches and re-stores its own stage into
        // the idle buffer (unconditional, so the loads stay in flight
        // across the matrix instructions instead of being copied at once)
        fetch(a0 + min(k + 1, nsub - 1) * kSkSub);
        if (ntw > 0) {
            const int col = 16 * kA;                // bytes: atoms 8 kA .. 8 kA + 7
            const int ra0 = (2 * (16 * w + rA) + cA) * kSkRowB + col;
            const int ra1 = (2 * (16 * w + 8 + rA) + cA) * kSkRowB + col;
            const half8_t Ah0 = *reinterpret_cast<const half8_t *>(cur + St::WH + ra0);
            const half8_t Al0 = *reinterpret_cast<const half8_t *>(cur + St::WL + ra0);
            const half8_t Ah1 = *reinterpret_cast<const half8_t *>(cur + St::WH + ra1);
            const half8_t Al1 = *reinterpret_cast<const half8_t *>(cur + St::WL + ra1);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t < ntw) {
                    const int rb = (2 * (8 * t + rA) + cA) * kSkRowB + col;
                    const half8_t Bh = *reinterpret_cast<const half8_t *>(cur + St::ZH + rb);
                    const half8_t Bl = *reinterpret_cast<const half8_t *>(cur + St::ZL + rb);
                    mfma_f4 p0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah0, Bh, zero4, 0, 0, 0);
                    mfma_f4 p1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah1, Bh, zero4, 0, 0, 0);
                    p0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah0, Bl, p0, 0, 0, 0);
                    p1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah1, Bl, p1, 0, 0, 0);
                    p0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al0, Bh, p0, 0, 0, 0);
                    p1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al1, Bh, p1, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        acc[0][t][v] += (double)p0[v];
                        acc[1][t][v] += (double)p1[v];
                    }
                }
            }
        }
        store(lds + ((k + 1) & 1) * St::BYTES);
        __syncthreads();
    }
    // results: through LDS so that one lane holds (Re U, Im U, Re V, Im V)
    // of one (row, j): M = 4 (l / 16) + v is (row r = M % 8, Re/Im = M / 8),
    // N = l % 16 is (j = N % 8, C/D = N / 8)
    double *scr = reinterpret_cast<double *>(lds) + w * 256;
    const int r = lane >> 3, jj = lane & 7;
    const int q0 = (r & 3) * 64 + 16 * (r >> 2) + jj;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t < ntw)
#pragma unroll
                for (int v = 0; v < 4; ++v) scr[v * 64 + lane] = acc[mt][t][v];
            __syncthreads();
            const int row = rw + 8 * mt + r;
            if (t < ntw && row < a.nfrows) {
                const double ur = scr[q0], ui = scr[q0 + 32], vr = scr[q0 + 8], vi = scr[q0 + 40];
                const FRow fr = ldv(gld(a.srows) + row);
                const int j = 8 * t + jj;
                if (j == 0) stv(gst(a.s0part) + (size_t)ch * a.nfrows + fr.pad0, make_double2(ur, ui));
                else if (j <= fr.m) stv(gst(a.ppart) + (size_t)ch * a.npk + fr.off + j - 1, make_double4(ur, ui, vr, vi));
            }
            __syncthreads();
        }
}

template <int NT>
__global__ __launch_bounds__(256, 2) void k_sk(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[2 * SkStage<NT>::BYTES];
    const int2 m = map[blockIdx.x];
    sk_body<NT>(args[m.x], m.y, lds);
}

template __global__ void k_sk<3>(const RecipArgs *, const int2 *);
template __global__ void k_sk<5>(const RecipArgs *, const int2 *);
template __global__ void k_sk<7>(const RecipArgs *, const int2 *);

// Energy and virial terms of one k vector (forces.f:1317-1331); returns
// A_k = 2 AKC S(k).
__device__ inline float2 k_terms(const RecipArgs &a, int ix, int iy, int iz, double sr, double si, double (&acc)[4]) {
    const double twopi = 2.0 * kPI;
    const double rxv = twopi * ix / a.box.l[0], ryv = twopi * iy / a.box.l[1], rzv = twopi * iz / a.box.l[2];
    double rk2 = rxv * rxv + ryv * ryv;
    rk2 = rk2 + rzv * rzv;
    const double alp2 = a.alpha * a.alpha;
    const double vol = a.box.l[0] * a.box.l[1] * a.box.l[2];
    const double cfur = 4.0 * kPI * a.coulf / vol;
    const double akc = cfur * exp(-0.25 * rk2 / alp2) / rk2;
    const double s2 = sr * sr + si * si;
    const double e = akc * s2;
    const double akv = 0.5 * akc * (4.0 * alp2 + rk2) / (alp2 * rk2);
    acc[0] += e;
    acc[1] += e - rxv * rxv * akv * s2;
    acc[2] += e - ryv * ryv * akv * s2;
    acc[3] += e - rzv * rzv * akv * s2;
    return make_float2((float)(2.0 * akc * sr), (float)(2.0 * akc * si));
}

// Sum the chunk partials in FP64 (fixed order), form S(+-j) = U_j +- i V_j,
// the energy QPE and virial projections of every k (forces.f:1317-1331)
// and the force-pass data: per row A of n_z = 0, and per pair entry, from
// D = (A+r - A-r, A+r + A-r, A+i + A-i, A+i - A-i), the coefficients of
// (Re p_j, Im p_j) in the four sums of recip_force_body,
//   S1: (-D3, D1)   S2: (D2, D4)   Z1: j (-D4, D2)   Z2: j (D1, D3),
// scaled by the tile's 2^e, split into f16 hi/lo and written in the operand
// order of its matrix instruction: step (j - 1) / 16; with u = (j - 1) mod
// 16, lane 16 (u / 4) + 4 row + s (s = 0..3 the sum), element 2 (u mod 4)
// + (0: Re p_j, 1: Im p_j); hi halves first, then lo.
__device__ inline void sk_reduce_body(const RecipArgs &a, int blk, double (*red)[256]) {
    // four lanes per item: lane q sums the chunks q, q + 4, ... and the
    // four lanes' sums are added in a fixed order (q 0 + 1, 2 + 3, then
    // both); for a pair entry lane q then writes sum q of the operand
    const int q = threadIdx.x & 3;
    const int e = blk * kSkRedItems + (threadIdx.x >> 2);
    auto quad_sum = [](double v) {
        v += __shfl_xor(v, 1);
        return v + __shfl_xor(v, 2);
    };
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (e < a.npk) {
        const int4 k = ldv(gld(a.fpk) + e);
        const FRow fr = ldv(gld(a.frows) + k.w);
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k.x >= 0 || k.y >= 0) {
            double ur = 0.0, ui = 0.0, vr = 0.0, vi = 0.0;
            for (int c = q; c < a.nchunk; c += 4) {
                const double4 p = ldv(gld(a.ppart) + (size_t)c * a.npk + e);
                ur += p.x;
                ui += p.y;
                vr += p.z;
                vi += p.w;
            }
            ur = quad_sum(ur);
            ui = quad_sum(ui);
            vr = quad_sum(vr);
            vi = quad_sum(vi);
            float2 ap = make_float2(0.f, 0.f), am = make_float2(0.f, 0.f);
            if (k.x >= 0) ap = k_terms(a, fr.ix, fr.iy, k.z, ur - vi, ui + vr, acc);
            if (k.y >= 0) am = k_terms(a, fr.ix, fr.iy, -k.z, ur + vi, ui - vr, acc);
            d = make_float4(ap.x - am.x, ap.x + am.x, ap.y + am.y, ap.y - am.y);
        }
        if (k.z <= fr.m) {
            const int jm = k.z - 1, u = jm & 15;
            const int rl = fr.pad2 & 255, e = (fr.pad2 >> 8) - 512;
            const float sc = ldexpf(1.f, e), fj = (float)k.z * sc;
            const float c[2][4] = {{-d.z * sc, d.y * sc, -fj * d.w, fj * d.x},
                                   {d.x * sc, d.w * sc, fj * d.y, fj * d.z}};
            MDX_GLOBAL _Float16 *w = reinterpret_cast<MDX_GLOBAL _Float16 *>(gst(a.wm)) + fr.pad1 +
                                     (size_t)(jm >> 4) * kForceStepH + 8 * (16 * (u >> 2) + 4 * rl) + 2 * (u & 3);
            for (int cm = 0; cm < 2; ++cm) {
                _Float16 hi, lo;
                split1(c[cm][q], hi, lo);
                w[8 * q + cm] = hi;
                w[512 + 8 * q + cm] = lo;
            }
        }
    } else if (e < a.npk + a.nfrows) {
        const int r = e - a.npk;
        const FRow fr = ldv(gld(a.frows) + r);
        float2 a0 = make_float2(0.f, 0.f);
        if (fr.k0 >= 0) {
            double sr = 0.0, si = 0.0;
            for (int c = q; c < a.nchunk; c += 4) {
                const double2 p = ldv(gld(a.s0part) + (size_t)c * a.nfrows + r);
                sr += p.x;
                si += p.y;
            }
            sr = quad_sum(sr);
            si = quad_sum(si);
            a0 = k_terms(a, fr.ix, fr.iy, 0, sr, si, acc);
        }
        if (q == 0) stv(gst(a.fa0) + r, a0);
    }
    if (q != 0)   // every lane of an item formed the same terms: count them once
        for (int i = 0; i < 4; ++i) acc[i] = 0.0;
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        MDX_GLOBAL double *o = gst(a.epart) + (size_t)blk * Q_N;
        for (int q = 0; q < Q_N; ++q) o[q] = 0.0;
        o[Q_PELS1] = red[0][0];
        o[Q_VIRX] = red[1][0];
        o[Q_VIRY] = red[2][0];
        o[Q_VIRZ] = red[3][0];
    }
}

__global__ __launch_bounds__(256) void k_sk_reduce(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double red[4][256];
    const int2 m = map[blockIdx.x];
    sk_reduce_body(args[m.x], m.y, red);
}

// Reciprocal forces (forces.f:1320-1326).  With b = q exp(2 pi i (n_x f_x +
// n_y f_y)) and p_j = exp(2 pi i j f_z), the two vectors n_z = +-j of a row
// contribute
//   t(+j) + t(-j) = b_r (p_i D1 - p_r D3) + b_i (p_r D2 + p_i D4)
//   t(+j) - t(-j) = b_r (p_i D2 - p_r D4) + b_i (p_r D1 + p_i D3)
// (t = Im(exp(i k.r) conj(A_k)), D from k_sk_reduce), so per atom and row
// the four sums over j
//   S1 = sum (p_i D1 - p_r D3)    S2 = sum (p_r D2 + p_i D4)
//   Z1 = sum j (p_i D2 - p_r D4)  Z2 = sum j (p_r D1 + p_i D3)
// are a matrix product (4 rows x 4 sums) x (16 j x (Re, Im)) x (16 atoms),
// v_mfma_f32_16x16x32_f16 on f16 hi/lo pairs (three products, FP32
// accumulation): the row coefficients are the A operand (laid out and
// scaled by 2^e per tile by k_sk_reduce), the atoms' p_j (FP64 recurrence,
// split once) the B operand, kept in registers for the whole row loop.
// Lane l then holds 2^e (S1, S2, Z1, Z2) of row l / 16 of the tile and atom
// l % 16, and adds
//   t = b_i A0_r - b_r A0_i + b_r S1 + b_i S2,  t_z = b_r Z1 + b_i Z2
// (A0: n_z = 0) into the FP64 sums F_x += n_x t, F_y += n_y t, F_z += t_z
// (t and t_z are FP32 terms of one atom and k row: their sums over n_z run
// inside the matrix instruction); b advances over the 4-row tiles of one
// n_x by exp(8 pi i f_y) (FP64) and restarts from an FP64 sincospi at
// every new n_x.  A wavefront takes
// kForceTiles tiles of 16 atoms, a block 4 wavefronts, each block one slice
// of the tiles; the 4 lanes of an atom are summed at the end and every
// slice writes its own partial force array (summed by the combine kernel).
template <int MZ>
__device__ inline void recip_force_body(const RecipArgs &a, int blk) {
    constexpr int S16 = MZ / 16;
    const int ab = blk % a.natom_blocks, sl = blk / a.natom_blocks;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int K = lane >> 4;
    const int t0 = sl * a.tiles_per_slice;
    const int t1 = min(t0 + a.tiles_per_slice, a.ntiles);
    const int abase = ab * kForceBlock + w * 16 * kForceTiles + (lane & 15);
    // B operand: element 2 q + c of step st is p_j, j = 16 st + 4 K + q + 1,
    // component c (Re, Im)
    half8_t Ph[kForceTiles][S16], Pl[kForceTiles][S16];
    double w4r[kForceTiles], w4i[kForceTiles], br[kForceTiles], bi[kForceTiles];
    double Fx[kForceTiles], Fy[kForceTiles], Fz[kForceTiles];
#pragma unroll
    for (int ta = 0; ta < kForceTiles; ++ta) {
        const int i = abase + 16 * ta;
        double4 at = make_double4(0, 0, 0, 0);
        double4 st = make_double4(1.0, 0.0, 1.0, 0.0);
        if (i < a.natoms) { at = ldv(gld(a.fq) + i); st = ldv(gld(a.stepz) + i); }
        float pv[S16][8];
        double wr = st.x, wi = st.y;
#pragma unroll
        for (int j = 1; j <= MZ; ++j) {
            const int u = (j - 1) & 15, stp = (j - 1) >> 4;
            if ((u >> 2) == K) {
                pv[stp][2 * (u & 3)] = (float)wr;
                pv[stp][2 * (u & 3) + 1] = (float)wi;
            }
            const double t = wr * st.x - wi * st.y;
            wi = wr * st.y + wi * st.x;
            wr = t;
        }
#pragma unroll
        for (int stp = 0; stp < S16; ++stp)
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                _Float16 hi, lo;
                split1(pv[stp][k], hi, lo);
                Ph[ta][stp][k] = hi;
                Pl[ta][stp][k] = lo;
            }
        double s4, c4;
        {
            double v = 4.0 * (at.y - rint(at.y));
            v -= rint(v);
            sincospi(2.0 * v, &s4, &c4);
        }
        w4r[ta] = c4;
        w4i[ta] = s4;
        br[ta] = bi[ta] = 0.0;
        Fx[ta] = Fy[ta] = Fz[ta] = 0.0;
    }
    const MDX_CONST FTile *tiles = sld(a.ftiles);
    const MDX_GLOBAL half8_t *wm = reinterpret_cast<const MDX_GLOBAL half8_t *>(gld(a.wm)) + lane;
    const MDX_GLOBAL float2 *fa0 = gld(a.fa0);
    int cix = -0x7fffffff, ciy = 0;
    // Two operand sets in ping-pong: the next tile's operand is requested
    // (unconditionally: past the slice end the last tile is re-read, and
    // the array ends with spare steps) while the current one computes.
    half8_t Ah0[S16], Al0[S16], Ah1[S16], Al1[S16];
    float2 a00 = make_float2(0.f, 0.f), a01 = make_float2(0.f, 0.f);
    FTile T0{}, T1{};
    auto load = [&](int tn, FTile &Tx, half8_t (&Ahx)[S16], half8_t (&Alx)[S16], float2 &a0x) {
        Tx = ldv(tiles + tn);
#pragma unroll
        for (int s = 0; s < S16; ++s) {
            Ahx[s] = wm[Tx.woff / 8 + s * (kForceStepH / 8)];
            Alx[s] = wm[Tx.woff / 8 + s * (kForceStepH / 8) + 64];
        }
        a0x = ldv(fa0 + Tx.r0 + K);
    };
    auto tile = [&](const FTile &T, const half8_t (&Ah)[S16], const half8_t (&Al)[S16], float2 a0raw) {
        const float2 a0 = K < T.nrow ? a0raw : make_float2(0.f, 0.f);
        const int ny = T.iy0 + K;
        const bool restart = T.ix != cix || T.iy0 != ciy + 4;
        cix = T.ix;
        ciy = T.iy0;
        mfma_f4 acc[kForceTiles];
#pragma unroll
        for (int ta = 0; ta < kForceTiles; ++ta) acc[ta] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S16; ++s) {
            if (s < T.nstep) {
#pragma unroll
                for (int ta = 0; ta < kForceTiles; ++ta) {
                    acc[ta] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[s], Ph[ta][s], acc[ta], 0, 0, 0);
                    acc[ta] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[s], Pl[ta][s], acc[ta], 0, 0, 0);
                    acc[ta] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[s], Ph[ta][s], acc[ta], 0, 0, 0);
                }
            }
        }
        if (restart) {
            // (f_x, f_y, q) again from memory: rare, and not worth registers
#pragma unroll
            for (int ta = 0; ta < kForceTiles; ++ta) {
                const int i = abase + 16 * ta;
                const double4 at = i < a.natoms ? ldv(gld(a.fq) + i) : make_double4(0, 0, 0, 0);
                double v = (double)T.ix * at.x + (double)ny * at.y;
                v = v - rint(v);
                double sv, cv;
                sincospi(2.0 * v, &sv, &cv);
                br[ta] = at.w * cv;
                bi[ta] = at.w * sv;
            }
        } else {
#pragma unroll
            for (int ta = 0; ta < kForceTiles; ++ta) {
                const double u = br[ta] * w4r[ta] - bi[ta] * w4i[ta];
                bi[ta] = br[ta] * w4i[ta] + bi[ta] * w4r[ta];
                br[ta] = u;
            }
        }
        const float isc = ldexpf(1.f, -T.pad0);
#pragma unroll
        for (int ta = 0; ta < kForceTiles; ++ta) {
            const float bre = (float)br[ta], bim = (float)bi[ta];
            const float ts = (bim * a0.x - bre * a0.y) + (bre * acc[ta][0] + bim * acc[ta][1]) * isc;
            const float tzs = (bre * acc[ta][2] + bim * acc[ta][3]) * isc;
            const double td = (double)ts;
            Fx[ta] = fma((double)T.ix, td, Fx[ta]);
            Fy[ta] = fma((double)ny, td, Fy[ta]);
            Fz[ta] += (double)tzs;
        }
    };
    if (t0 < t1) load(t0, T0, Ah0, Al0, a00);
    for (int t = t0; t < t1; t += 2) {
        load(min(t + 1, t1 - 1), T1, Ah1, Al1, a01);
        tile(T0, Ah0, Al0, a00);
        if (t + 1 >= t1) break;
        load(min(t + 2, t1 - 1), T0, Ah0, Al0, a00);
        tile(T1, Ah1, Al1, a01);
    }
    const double tx = 2.0 * kPI / a.box.l[0], ty = 2.0 * kPI / a.box.l[1], tz = 2.0 * kPI / a.box.l[2];
    MDX_GLOBAL double *o = gst(a.rpart) + (size_t)sl * 3 * a.natoms;
#pragma unroll
    for (int ta = 0; ta < kForceTiles; ++ta) {
        // the atom's four lanes (rows 0..3 of every tile), fixed order
        double x = Fx[ta], y = Fy[ta], z = Fz[ta];
        x += __shfl_xor(x, 16); y += __shfl_xor(y, 16); z += __shfl_xor(z, 16);
        x += __shfl_xor(x, 32); y += __shfl_xor(y, 32); z += __shfl_xor(z, 32);
        const int i = abase + 16 * ta;
        if (K == 0 && i < a.natoms) {
            o[i] = tx * x;
            o[a.natoms + i] = ty * y;
            o[2 * a.natoms + i] = tz * z;
        }
    }
}

// Three wavefronts per SIMD up to 32 values of n_z per row (three 16-atom
// tiles per wavefront, 159 VGPRs: 6 % faster than four tiles at two
// wavefronts per SIMD at 96,174 atoms); 48 values need the registers of two.
template <int MZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MZ <= 32 ? 3 : 2)))
void k_recip_force(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    recip_force_body<MZ>(args[m.x], m.y);
}

template __global__ void k_recip_force<16>(const RecipArgs *, const int2 *);
template __global__ void k_recip_force<32>(const RecipArgs *, const int2 *);
template __global__ void k_recip_force<48>(const RecipArgs *, const int2 *);

}  // namespace mdx
