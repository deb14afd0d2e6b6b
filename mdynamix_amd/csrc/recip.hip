// Ewald reciprocal sum (replaces FURIR, forces.f:1217-1339).
//
// The half-space k set of INIFUR (forces.f:1356-1427) is regrouped into
// rows of fixed (n_x, n_y) whose n_z run over a contiguous range; rows are
// sorted by length so the 64 rows of one wavefront have similar lengths.
//
// k_sk: lane = row, atoms of one chunk streamed with uniform (scalar) loads.
//   For each atom the phase 2 pi (n_x f_x + n_y f_y + n_z0 f_z) is formed
//   in FP64 fractional coordinates and reduced to one turn before sincospi;
//   the remaining terms of the row follow by the recurrence
//   P_{j+1} = 2 cos(theta) P_j - P_{j-1}, theta = 2 pi f_z, re-anchored in
//   FP64 every 8 terms.  Per-chunk FP32 sums are written as
//   partials and added in FP64 by k_sk_reduce, which also forms the energy
//   QPE, the virial projections and A_k = 2 AKC S(k) for the force pass.
// k_recip_force: lane = atom, rows streamed uniformly; the force on an atom
//   is q sum_k k Im(exp(i k.r) conj(A_k)) (forces.f:1320-1326), evaluated
//   for the k vectors +-n_z together on a per-atom table of exp(2 pi i j f_z)
//   (see recip_force_body).
#include "mdx_internal.h"

namespace mdx {

__device__ inline void recip_prep_body(const RecipArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.natoms) return;
    const double fx = a.x[i] / a.box.l[0], fy = a.y[i] / a.box.l[1], fz = a.z[i] / a.box.l[2];
    a.fq[i] = make_double4(fx, fy, fz, a.q[i]);
    // exp(2 pi i f_z) and its 8th power, in FP64
    const double u = fz - rint(fz);
    double s1, c1, s8, c8;
    sincospi(2.0 * u, &s1, &c1);
    double u8 = 8.0 * u;
    u8 -= rint(u8);
    sincospi(2.0 * u8, &s8, &c8);
    a.stepz[i] = make_double4(c1, s1, c8, s8);
}

__global__ __launch_bounds__(256) void k_recip_prep(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    recip_prep_body(args[m.x], m.y);
}

// exp(i 2 pi u) at the start of a row, u reduced to one turn in FP64.
__device__ inline void row_start(const RecipRow &r, double4 a, double &c, double &s) {
    double u = (double)r.nx * a.x + (double)r.ny * a.y;
    u = u + (double)r.nz0 * a.z;
    u = u - rint(u);
    float sf, cf;
    sincospif((float)(2.0 * u), &sf, &cf);
    c = (double)cf;
    s = (double)sf;
}

// Structure factors of 64 rows over one atom chunk.  Along a row the
// terms P_j = q exp(i(phi0 + j theta)) follow the three-term recurrence
// P_{j+1} = 2 cos(theta) P_j - P_{j-1} (2 FMA per term), in blocks of 8
// seeded from an FP64 anchor advanced by exp(8 i theta), so rounding
// cannot build up over more than 7 steps.  The atom data of the chunk is
// wavefront-uniform and read with scalar loads, one atom ahead.
template <int MAXLEN>
__device__ inline void sk_body(const RecipArgs &a, int blk) {
    const int rg = blk % a.nrowgroups, ch = blk / a.nrowgroups;
    const int r = rg * 64 + threadIdx.x;
    const int a0 = ch * a.chunk;
    const int a1 = min(a0 + a.chunk, a.natoms);
    const int wlen = sld(a.wavelen)[rg];          // longest row of this wavefront
    const MDX_CONST double4 *fq = sld(a.fq);
    const MDX_CONST double4 *stz = sld(a.stepz);
    RecipRow row = {0, 0, 0, 0};
    if (r < a.nrows) row = ldv(gld(a.rows) + r);
    float are[MAXLEN], aim[MAXLEN];
#pragma unroll
    for (int j = 0; j < MAXLEN; ++j) { are[j] = 0.f; aim[j] = 0.f; }
    double4 at = ldv(fq + a0), st = ldv(stz + a0);
    for (int i = a0; i < a1; ++i) {
        const double4 atc = at, stc = st;
        if (i + 1 < a1) { at = ldv(fq + i + 1); st = ldv(stz + i + 1); }
        double s0, c0;
        row_start(row, atc, c0, s0);
        double Ar = atc.w * c0, Ai = atc.w * s0;           // anchor, FP64
        const float sr = (float)stc.x, si = (float)stc.y, c2 = 2.0f * sr;
#pragma unroll
        for (int j0 = 0; j0 < MAXLEN; j0 += 8) {
            if (j0 >= wlen) break;
            float pr0 = (float)Ar, pi0 = (float)Ai;
            float pr1 = pr0 * sr - pi0 * si, pi1 = pr0 * si + pi0 * sr;
            are[j0] += pr0; aim[j0] += pi0;
            are[j0 + 1] += pr1; aim[j0 + 1] += pi1;
#pragma unroll
            for (int j = j0 + 2; j < j0 + 8; ++j) {
                const float pr2 = fmaf(c2, pr1, -pr0), pi2 = fmaf(c2, pi1, -pi0);
                are[j] += pr2; aim[j] += pi2;
                pr0 = pr1; pi0 = pi1; pr1 = pr2; pi1 = pi2;
            }
            const double tr = Ar * stc.z - Ai * stc.w;
            Ai = Ar * stc.w + Ai * stc.z;
            Ar = tr;
        }
    }
    if (r < a.nrows) {
        const int off = gld(a.rowoff)[r];
        MDX_GLOBAL float2 *o = gst(a.part) + (size_t)ch * a.nkv + off;
#pragma unroll
        for (int j = 0; j < MAXLEN; ++j)
            if (j < row.len) stv(o + j, make_float2(are[j], aim[j]));
    }
}

template <int MAXLEN>
__global__ __launch_bounds__(64) void k_sk(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    sk_body<MAXLEN>(args[m.x], m.y);
}

template __global__ void k_sk<16>(const RecipArgs *, const int2 *);
template __global__ void k_sk<32>(const RecipArgs *, const int2 *);
template __global__ void k_sk<64>(const RecipArgs *, const int2 *);
template __global__ void k_sk<96>(const RecipArgs *, const int2 *);

// Sum chunk partials in FP64 (fixed order), energy and virial terms of each
// k (forces.f:1317-1331) and A_k = 2 AKC S(k).  kidx holds (nx, ny, nz) of
// every k in row order.
__device__ inline void sk_reduce_body(const RecipArgs &a, int blk, double (*red)[256]) {
    const int k = blk * 256 + threadIdx.x;
    double e = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (k < a.nkv) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < a.nchunk; ++c) {
            const float2 p = a.part[(size_t)c * a.nkv + k];
            sr += (double)p.x;
            si += (double)p.y;
        }
        const int4 id = a.kidx[k];
        const double twopi = 2.0 * kPI;
        const double rxv = twopi * id.x / a.box.l[0], ryv = twopi * id.y / a.box.l[1], rzv = twopi * id.z / a.box.l[2];
        double rk2 = rxv * rxv + ryv * ryv;
        rk2 = rk2 + rzv * rzv;
        const double alp2 = a.alpha * a.alpha;
        const double vol = a.box.l[0] * a.box.l[1] * a.box.l[2];
        const double cfur = 4.0 * kPI * a.coulf / vol;
        const double akc = cfur * exp(-0.25 * rk2 / alp2) / rk2;
        const double s2 = sr * sr + si * si;
        e = akc * s2;
        const double akv = 0.5 * akc * (4.0 * alp2 + rk2) / (alp2 * rk2);
        vx = e - rxv * rxv * akv * s2;
        vy = e - ryv * ryv * akv * s2;
        vz = e - rzv * rzv * akv * s2;
        a.ak[k] = make_float2((float)(2.0 * akc * sr), (float)(2.0 * akc * si));
        a.sk[k] = make_double2(sr, si);
    }
    red[0][threadIdx.x] = e; red[1][threadIdx.x] = vx; red[2][threadIdx.x] = vy; red[3][threadIdx.x] = vz;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = a.epart + (size_t)blk * Q_N;
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

// Pair packing for the force pass: for each (row, j) the sums and
// differences of A_k = 2 AKC S(k) at n_z = +j and -j; per row the n_z = 0
// value.
__device__ inline void fpack_body(const RecipArgs &a, int blk) {
    const int e = blk * 256 + threadIdx.x;
    if (e < a.npk) {
        const int4 k = ldv(gld(a.fpk) + e);
        const float2 ap = k.x >= 0 ? ldv(gld(a.ak) + k.x) : make_float2(0.f, 0.f);
        const float2 am = k.y >= 0 ? ldv(gld(a.ak) + k.y) : make_float2(0.f, 0.f);
        stv(gst(a.fpack) + e, make_float4(ap.x - am.x, ap.x + am.x, ap.y + am.y, ap.y - am.y));
    } else if (e < a.npk + a.nfrows) {
        const int r = e - a.npk;
        const int k0 = gld(a.frows)[r].k0;
        stv(gst(a.fa0) + r, k0 >= 0 ? ldv(gld(a.ak) + k0) : make_float2(0.f, 0.f));
    }
}

__global__ __launch_bounds__(256) void k_fpack(const RecipArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    fpack_body(args[m.x], m.y);
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
