// Ewald reciprocal sum (replaces FURIR, forces.f:1217-1339).
//
// The half-space k set of INIFUR (forces.f:1356-1427) is regrouped into
// rows of fixed (n_x, n_y); within a row the vectors n_z = +j and -j are
// taken together (the k set is symmetric in n_z except the (0, 0) row).
// k_recip_prep: per-atom phase tables along x, y and z (FP64 recurrences);
// k_sk: S(+-j) = U_j +- i V_j, with U_j, V_j sums of B p_j over the atoms
//   of a chunk (B = q exp(2 pi i (n_x f_x + n_y f_y)), p_j = exp(2 pi i j
//   f_z)), as a matrix product on the matrix cores;
// k_sk_reduce: FP64 sums of the chunk partials, energy, virial, and the
//   per-pair data of the force pass;
// k_recip_force: lane = atom, rows uniform, the force
//   q sum_k k Im(exp(i k.r) conj(A_k)) (forces.f:1320-1326) evaluated for
//   +-j together (see recip_force_body).
#include "mdx_internal.h"

namespace mdx {

__device__ inline void recip_prep_body(const RecipArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.natoms) return;
    const double fx = a.x[i] / a.box.l[0], fy = a.y[i] / a.box.l[1], fz = a.z[i] / a.box.l[2];
    const double q = a.q[i];
    stv(gst(a.fq) + i, make_double4(fx, fy, fz, q));
    // exp(2 pi i f_z) and its 8th power, in FP64
    const double u = fz - rint(fz);
    double s1, c1, s8, c8;
    sincospi(2.0 * u, &s1, &c1);
    double u8 = 8.0 * u;
    u8 -= rint(u8);
    sincospi(2.0 * u8, &s8, &c8);
    stv(gst(a.stepz) + i, make_double4(c1, s1, c8, s8));
    // phase tables (FP64 recurrences from FP64 sincospi steps, FP32 values)
    const size_t n = (size_t)a.natoms;
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
    {
        double pr = 1.0, pi = 0.0;
        MDX_GLOBAL float2 *t = gst(a.tz) + i;
        for (int k = 0; k < a.mzp; ++k) {
            stv(t + k * n, make_float2((float)pr, (float)pi));
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

// Structure factors on the matrix cores.  For a row (n_x, n_y) with
// B_a = q_a exp(2 pi i (n_x f_x + n_y f_y)) and p_j = exp(2 pi i j f_z) =
// C_j + i D_j, the vectors n_z = +-j have S(+-j) = U_j +- i V_j with
// U_j = sum_a B_a C_j(a), V_j = sum_a B_a D_j(a); n_z = 0 is U_0 (C_0 = 1).
// Over the atoms of a chunk this is the real product
//   [Re B; Im B] (rows x atoms)  x  [C_j, D_j] (atoms x j)
// evaluated with v_mfma_f32_16x16x4_f32: an M tile is 8 rows x (Re, Im),
// an N tile 8 values of j x (C, D), K = 4 atoms.  A block holds 64 rows
// (sorted by pair count; a wavefront takes 16 rows = 2 M tiles and as many
// N tiles as its first, longest row needs) and one chunk of atoms, staged
// through LDS 32 atoms at a time: B = tx[n_x] ty[|n_y|] (conjugated for
// n_y < 0) and p_j = tz[j].  Operand layout of the instruction (lane l):
// A[l % 16][l / 16], B[l / 16][l % 16]; result D[4 (l / 16) + v][l % 16].
// Each block writes its FP32 chunk sums per pair entry (U_j, V_j) and per
// row (S(n_z = 0)); k_sk_reduce adds the chunks in FP64.
// Blocks are numbered so that the row groups of one chunk share an XCD
// (block index mod 8), whose L2 then serves that chunk's table columns.
constexpr int kSkStride = kSkSub + 4;      // float2 per LDS row (bank spread)

template <int NT>
__device__ inline void sk_body(const RecipArgs &a, int blk, float2 *wl, float2 *zl) {
    const int G = a.nsgroups;
    const int x8 = blk & 7, s8 = blk >> 3;
    const int g = s8 % G, ch = (s8 / G) * 8 + x8;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int a0 = ch * a.chunk;
    const int a1 = min(a0 + a.chunk, a.natoms);
    const int rbase = g * kSkRows;
    const size_t n = (size_t)a.natoms;
    // staging assignment: atom tid % 32 of the stage, rows / j values tid / 32 + 8 i
    const int sa = tid & 31, sr0 = tid >> 5;
    int offx[8], offy[8];
    bool conj[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int r = rbase + sr0 + 8 * i;
        offx[i] = -1;
        offy[i] = 0;
        conj[i] = false;
        if (r < a.nfrows) {
            const FRow fr = ldv(gld(a.srows) + r);
            offx[i] = fr.ix * a.natoms;
            offy[i] = abs(fr.iy) * a.natoms;
            conj[i] = fr.iy < 0;
        }
    }
    // N tiles: the group's first row is its longest, a wavefront's first row is its longest
    const int rw = rbase + 16 * w;
    const int ntw = rw < a.nfrows ? min(NT, (ldv(gld(a.srows) + rw).m + 8) >> 3) : 0;
    const int ntg = rbase < a.nfrows ? min(NT, (ldv(gld(a.srows) + rbase).m + 8) >> 3) : 0;
    mfma_f4 acc[2][NT];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    const MDX_GLOBAL float2 *tx = gld(a.tx), *ty = gld(a.ty), *tz = gld(a.tz);
    const int rA = lane & 7, cA = (lane >> 3) & 1, kA = lane >> 4;
    const float *wlf = reinterpret_cast<const float *>(wl);
    const float *zlf = reinterpret_cast<const float *>(zl);
    for (int sb = a0; sb < a1; sb += kSkSub) {
        const int ia = sb + sa;
        const bool ok = ia < a1;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            float2 b = make_float2(0.f, 0.f);
            if (ok && offx[i] >= 0) {
                const float2 X = ldv(tx + offx[i] + ia);
                float2 Y = ldv(ty + offy[i] + ia);
                if (conj[i]) Y.y = -Y.y;
                b = make_float2(X.x * Y.x - X.y * Y.y, X.x * Y.y + X.y * Y.x);
            }
            wl[(sr0 + 8 * i) * kSkStride + sa] = b;
        }
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            const int j = sr0 + 8 * i;
            float2 p = make_float2(0.f, 0.f);
            if (ok && i < ntg) p = ldv(tz + (size_t)j * n + ia);
            zl[j * kSkStride + sa] = p;
        }
        __syncthreads();
        if (ntw > 0) {
#pragma unroll
            for (int s = 0; s < kSkSub / 4; ++s) {
                const int at = 4 * s + kA;
                const float A0 = wlf[2 * ((16 * w + rA) * kSkStride + at) + cA];
                const float A1 = wlf[2 * ((16 * w + 8 + rA) * kSkStride + at) + cA];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    if (t < ntw) {
                        const float Bv = zlf[2 * ((8 * t + rA) * kSkStride + at) + cA];
                        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A0, Bv, acc[0][t], 0, 0, 0);
                        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(A1, Bv, acc[1][t], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
    // results: through LDS so that one lane holds (Re U, Im U, Re V, Im V)
    // of one (row, j): M = 4 (l / 16) + v is (row r = M % 8, Re/Im = M / 8),
    // N = l % 16 is (j = N % 8, C/D = N / 8)
    float *scr = reinterpret_cast<float *>(wl) + w * 256;
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
                const float ur = scr[q0], ui = scr[q0 + 32], vr = scr[q0 + 8], vi = scr[q0 + 40];
                const FRow fr = ldv(gld(a.srows) + row);
                const int j = 8 * t + jj;
                if (j == 0) stv(gst(a.s0part) + (size_t)ch * a.nfrows + fr.pad0, make_float2(ur, ui));
                else if (j <= fr.m) stv(gst(a.ppart) + (size_t)ch * a.npk + fr.off + j - 1, make_float4(ur, ui, vr, vi));
            }
            __syncthreads();
        }
}

template <int NT>
__global__ __launch_bounds__(256, 2) void k_sk(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ float2 wl[kSkRows * kSkStride];
    __shared__ float2 zl[NT * 8 * kSkStride];
    const int2 m = map[blockIdx.x];
    sk_body<NT>(args[m.x], m.y, wl, zl);
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
// and the force-pass data: per pair entry D = (A+r - A-r, A+r + A-r,
// A+i + A-i, A+i - A-i), per row A of n_z = 0.
__device__ inline void sk_reduce_body(const RecipArgs &a, int blk, double (*red)[256]) {
    const int e = blk * 256 + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (e < a.npk) {
        const int4 k = ldv(gld(a.fpk) + e);
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k.x >= 0 || k.y >= 0) {
            double ur = 0.0, ui = 0.0, vr = 0.0, vi = 0.0;
            for (int c = 0; c < a.nchunk; ++c) {
                const float4 p = ldv(gld(a.ppart) + (size_t)c * a.npk + e);
                ur += (double)p.x;
                ui += (double)p.y;
                vr += (double)p.z;
                vi += (double)p.w;
            }
            const FRow fr = ldv(gld(a.frows) + k.w);
            float2 ap = make_float2(0.f, 0.f), am = make_float2(0.f, 0.f);
            if (k.x >= 0) ap = k_terms(a, fr.ix, fr.iy, k.z, ur - vi, ui + vr, acc);
            if (k.y >= 0) am = k_terms(a, fr.ix, fr.iy, -k.z, ur + vi, ui - vr, acc);
            d = make_float4(ap.x - am.x, ap.x + am.x, ap.y + am.y, ap.y - am.y);
        }
        stv(gst(a.fpack) + e, d);
    } else if (e < a.npk + a.nfrows) {
        const int r = e - a.npk;
        const FRow fr = ldv(gld(a.frows) + r);
        float2 a0 = make_float2(0.f, 0.f);
        if (fr.k0 >= 0) {
            double sr = 0.0, si = 0.0;
            for (int c = 0; c < a.nchunk; ++c) {
                const float2 p = ldv(gld(a.s0part) + (size_t)c * a.nfrows + r);
                sr += (double)p.x;
                si += (double)p.y;
            }
            a0 = k_terms(a, fr.ix, fr.iy, 0, sr, si, acc);
        }
        stv(gst(a.fa0) + r, a0);
    }
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

// Reciprocal forces (forces.f:1320-1326); a block covers kForceBlock atoms
// (four wavefronts on one CU, which share the row data through the scalar
// cache) and one slice of rows; each slice writes its own partial force
// array (summed by the combine kernel).  With b = q exp(2 pi i (n_x f_x + n_y f_y)) and
// p_j = exp(2 pi i j f_z), the two vectors n_z = +-j contribute
//   t(+j) + t(-j) = b_r (p_i D1 - p_r D3) + b_i (p_r D2 + p_i D4)
//   t(+j) - t(-j) = b_r (p_i D2 - p_r D4) + b_i (p_r D1 + p_i D3)
// with t = Im(exp(i k.r) conj(A_k)), so a row costs 10 VALU operations per
// pair of k vectors on the per-atom table p_j (FP64 recurrence, kept in
// registers) and the row's pair data (wavefront-uniform, scalar loads).  b advances
// along n_y by exp(2 pi i f_y) in FP64 and restarts from an FP64 sincospi
// at every new n_x.
template <int MZ>
__device__ inline void recip_force_body(const RecipArgs &a, int blk) {
    const int ab = blk % a.natom_blocks, sl = blk / a.natom_blocks;
    const int i = ab * kForceBlock + threadIdx.x;
    const int r0 = sl * a.rows_per_slice;
    const int r1 = min(r0 + a.rows_per_slice, a.nfrows);
    double4 at = make_double4(0, 0, 0, 0);
    double4 st = make_double4(1.0, 0.0, 1.0, 0.0);
    if (i < a.natoms) { at = ldv(gld(a.fq) + i); st = ldv(gld(a.stepz) + i); }
    float pr[MZ], pim[MZ];
    {
        double wr = st.x, wi = st.y;
#pragma unroll
        for (int j = 0; j < MZ; ++j) {
            pr[j] = (float)wr;
            pim[j] = (float)wi;
            const double t = wr * st.x - wi * st.y;
            wi = wr * st.y + wi * st.x;
            wr = t;
        }
    }
    double eyr, eyi;
    {
        const double v = at.y - rint(at.y);
        sincospi(2.0 * v, &eyi, &eyr);
    }
    const double tx = 2.0 * kPI / a.box.l[0], ty = 2.0 * kPI / a.box.l[1], tz = 2.0 * kPI / a.box.l[2];
    const MDX_CONST FRow *frows = sld(a.frows);
    const MDX_CONST float4 *fpack = sld(a.fpack);
    const MDX_CONST float2 *fa0 = sld(a.fa0);
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    if (r0 < r1) {
        double Br = 0.0, Bi = 0.0;
        int cix = -0x7fffffff, ciy = 0;
        // the pair data of consecutive rows is contiguous: the blocks after a
        // row's last one are the next row's first ones (the array ends with
        // two spare blocks), so blocks are requested two blocks ahead into a
        // three-block ring of registers whose roles rotate at compile time
        FRow nfr = ldv(frows + r0);
        float4 buf[3][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            buf[0][q] = ldv(fpack + nfr.off + q);
            buf[1][q] = ldv(fpack + nfr.off + 4 + q);
        }
        for (int r = r0; r < r1; ++r) {
            const FRow fr = nfr;
            if (r + 1 < r1) nfr = ldv(frows + r + 1);
            if (fr.ix != cix || fr.iy != ciy + 1) {
                double v = (double)fr.ix * at.x + (double)fr.iy * at.y;
                v = v - rint(v);
                double s, c;
                sincospi(2.0 * v, &s, &c);
                Br = at.w * c;
                Bi = at.w * s;
            } else {
                const double t = Br * eyr - Bi * eyi;
                Bi = Br * eyi + Bi * eyr;
                Br = t;
            }
            cix = fr.ix;
            ciy = fr.iy;
            const float2 a0 = ldv(fa0 + r);
            const MDX_CONST float4 *P = fpack + fr.off;
            const int nb = (fr.m + 3) >> 2;
            float sx = 0.f, sy = 0.f, zx = 0.f, zy = 0.f;
#pragma unroll
            for (int jb = 0; jb < MZ / 4; ++jb) {
                if (jb >= nb) break;
#pragma unroll
                for (int q = 0; q < 4; ++q) buf[(jb + 2) % 3][q] = ldv(P + (jb + 2) * 4 + q);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = jb * 4 + q;
                    const float4 d = buf[jb % 3][q];
                    sx = fmaf(pim[j], d.x, sx);
                    sx = fmaf(-pr[j], d.z, sx);
                    sy = fmaf(pr[j], d.y, sy);
                    sy = fmaf(pim[j], d.w, sy);
                    // n_z-weighted: (j + 1) (p_i D2 - p_r D4), (j + 1) (p_r D1 + p_i D3)
                    const float ux = fmaf(pim[j], d.y, -pr[j] * d.w);
                    const float uy = fmaf(pr[j], d.x, pim[j] * d.z);
                    zx = fmaf((float)(j + 1), ux, zx);
                    zy = fmaf((float)(j + 1), uy, zy);
                }
            }
            // the next row's first two blocks sit in buf[nb % 3], buf[(nb + 1) % 3]
            const int rot = nb % 3;
            if (rot == 1) {
#pragma unroll
                for (int q = 0; q < 4; ++q) { buf[0][q] = buf[1][q]; buf[1][q] = buf[2][q]; }
            } else if (rot == 2) {
#pragma unroll
                for (int q = 0; q < 4; ++q) { const float4 t = buf[0][q]; buf[0][q] = buf[2][q]; buf[1][q] = t; }
            }
            const float br = (float)Br, bi = (float)Bi;
            const float ts = (bi * a0.x - br * a0.y) + (br * sx + bi * sy);
            const float tzs = br * zx + bi * zy;
            Fx += tx * fr.ix * (double)ts;
            Fy += ty * fr.iy * (double)ts;
            Fz += tz * (double)tzs;
        }
    }
    if (i < a.natoms) {
        MDX_GLOBAL double *o = gst(a.rpart) + (size_t)sl * 3 * a.natoms;
        o[i] = Fx;
        o[a.natoms + i] = Fy;
        o[2 * a.natoms + i] = Fz;
    }
}

template <int MZ>
__global__ __launch_bounds__(kForceBlock, 4) void k_recip_force(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    recip_force_body<MZ>(args[m.x], m.y);
}

template __global__ void k_recip_force<16>(const RecipArgs *, const int2 *);
template __global__ void k_recip_force<32>(const RecipArgs *, const int2 *);
template __global__ void k_recip_force<48>(const RecipArgs *, const int2 *);

}  // namespace mdx
