// Ewald reciprocal sum in FP64 on device arrays (FURIR, forces.f:1217-1339,
// with the k set of INIFUR, forces.f:1356-1427), stateless: no engine
// handle, every buffer in one caller workspace.  It is the FP64 reference
// path of the reciprocal sum on the GPU (the engine's default path forms
// the same sums on the matrix cores from FP32 phase tables) and the
// reciprocal part of the engine's FP64 precision mode.
//
// Workspace (FurLayout): the k vectors as (n_x, n_y, n_z) in INIFUR order,
// the structure factor's partial sums per (atom chunk, k), A_k = 2 AKC S(k)
// per k, and per-block energy/virial partials.  Every entry is written
// before it is read, so nothing needs clearing between calls, and every sum
// runs in a fixed order (bit-identical results call after call).
//
//   k_fur_kset    one block: the k set in INIFUR order (ballot compaction
//                 over the cube of candidate vectors, cells in loop order)
//   k_fur_sk      thread = k, block = (256 k, atom chunk): partial S(k)
//                 = sum q (cos, sin)(k.r) over the chunk's atoms (LDS-staged)
//   k_fur_sred    thread = k: S(k) from the chunk partials (chunk order),
//                 A_k, energy AKC |S|^2 and the virial projections
//   k_fur_force   thread = atom: F = sum_k k q (Re A sin - Im A cos)(k.r)
//                 over the k set staged through LDS (k order)
//   k_fur_esum    one block: energy and virial totals (block order)
// Phases are 2 pi (n . f) with fractional coordinates f = r / L, reduced
// to [-1/2, 1/2] turns before sincospi.  Inside FURIR's 5 % window of box
// changes the k set stays the one enumerated at the old box while the
// vectors 2 pi n / L take the current one (forces.f:1250-1276): the
// enumeration box (klx, kly, klz) is given apart from the current box.
#include "mdx_internal.h"

namespace mdx {

constexpr int kFurChunk = 2048;    // atoms per structure-factor block
constexpr int kFurT = 256;         // threads per block of the FURIR kernels

struct FurArgs {
    int n, nkv, nchunk;
    int kmax, kmay, kmaz;          // cube of candidate vectors (INIFUR KMAX..)
    double cutk2;                  // 4 alpha^2 FEXP
    double lx, ly, lz, alp2, cfur; // box, alpha^2, 4 pi COULF / V
    double klx, kly, klz;          // box of the k-set enumeration (INIFUR)
    const double *x, *y, *z, *q;
    double *fx, *fy, *fz;          // forces out
    double *e4;                    // QPE, VIRX, VIRY, VIRZ out (or null)
    double *qrow;                  // or null: an engine energy row (Q_N slots) to fill
    int4 *kv;                      // [nkv]
    double2 *spart;                // [nchunk][nkv]
    double2 *ak;                   // [nkv]
    double4 *epart;                // [ceil(nkv / kFurT)]
    int *count;                    // enumerated count (checked against nkv)
};

// k vectors with k^2 <= CUTK2 in INIFUR's loop order: IKZ = -KMAZ..KMAZ,
// IKX = 0..KMAX, IKY from -KMAY (IKX > 0), 0 (IKX = 0, IKZ > 0) or 1
// (IKX = 0, IKZ <= 0) to KMAY; RK2 = ((2 pi IKX / L)^2 + (2 pi IKY / L)^2)
// + (2 pi IKZ / L)^2 as INIFUR evaluates it.
__global__ __launch_bounds__(1024) void k_fur_kset(FurArgs a) {
#pragma clang fp contract(off)
    __shared__ int wsum[16];
    __shared__ int base_s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ny = 2 * a.kmay + 1, nx = a.kmax + 1, nz = 2 * a.kmaz + 1;
    const long long ncell = (long long)nz * nx * ny;
    const double twopi = 2.0 * kPI;
    if (tid == 0) base_s = 0;
    __syncthreads();
    for (long long c0 = 0; c0 < ncell; c0 += 1024) {
        const long long c = c0 + tid;
        bool keep = false;
        int ix = 0, iy = 0, iz = 0;
        if (c < ncell) {
            iz = (int)(c / ((long long)nx * ny)) - a.kmaz;
            const int rem = (int)(c % ((long long)nx * ny));
            ix = rem / ny;
            iy = rem % ny - a.kmay;
            const bool half = ix > 0 || iy > 0 || (iy == 0 && iz > 0);
            if (half) {
                const double rxv = twopi * ix, ryv = twopi * iy, rzv = twopi * iz;
                const double bx = rxv / a.klx, by = ryv / a.kly, bz = rzv / a.klz;
                double rk2 = bx * bx + by * by;
                rk2 = rk2 + bz * bz;
                keep = rk2 <= a.cutk2;
            }
        }
        const unsigned long long m = __ballot(keep);
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int before = base_s;
        for (int k = 0; k < w; ++k) before += wsum[k];
        const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const int pos = before + __popcll(m & below);
        if (keep && pos < a.nkv) a.kv[pos] = make_int4(ix, iy, iz, 0);
        __syncthreads();
        if (tid == 0)
            for (int k = 0; k < 16; ++k) base_s += wsum[k];
        __syncthreads();
    }
    if (tid == 0) *a.count = base_s;
}

// Phase turns n . f of a k vector and an atom, reduced to [-1/2, 1/2].
__device__ inline double fur_turns(const int4 &k, double fx, double fy, double fz) {
    double v = (double)k.x * fx + (double)k.y * fy;
    v = v + (double)k.z * fz;
    return v - rint(v);
}

__global__ __launch_bounds__(kFurT) void k_fur_sk(FurArgs a) {
    __shared__ double sx[kFurT], sy[kFurT], sz[kFurT], sq[kFurT];
    const int kk = blockIdx.x * kFurT + threadIdx.x;
    const int ch = blockIdx.y;
    const int a0 = ch * kFurChunk, a1 = min(a0 + kFurChunk, a.n);
    const int4 k = kk < a.nkv ? a.kv[kk] : make_int4(0, 0, 0, 0);
    double sr = 0.0, si = 0.0;
    for (int b = a0; b < a1; b += kFurT) {
        const int i = b + threadIdx.x;
        if (i < a1) {
            sx[threadIdx.x] = a.x[i] / a.lx;
            sy[threadIdx.x] = a.y[i] / a.ly;
            sz[threadIdx.x] = a.z[i] / a.lz;
            sq[threadIdx.x] = a.q[i];
        }
        __syncthreads();
        const int m = min(kFurT, a1 - b);
        for (int t = 0; t < m; ++t) {
            double s, c;
            sincospi(2.0 * fur_turns(k, sx[t], sy[t], sz[t]), &s, &c);
            sr += sq[t] * c;
            si += sq[t] * s;
        }
        __syncthreads();
    }
    if (kk < a.nkv) a.spart[(size_t)ch * a.nkv + kk] = make_double2(sr, si);
}

__global__ __launch_bounds__(kFurT) void k_fur_sred(FurArgs a) {
    __shared__ double red[4][kFurT];
    const int kk = blockIdx.x * kFurT + threadIdx.x;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    if (kk < a.nkv) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < a.nchunk; ++c) {
            const double2 p = a.spart[(size_t)c * a.nkv + kk];
            sr += p.x;
            si += p.y;
        }
        const int4 k = a.kv[kk];
        const double twopi = 2.0 * kPI;
        const double rxv = twopi * k.x / a.lx, ryv = twopi * k.y / a.ly, rzv = twopi * k.z / a.lz;
        double rk2 = rxv * rxv + ryv * ryv;
        rk2 = rk2 + rzv * rzv;
        const double akc = a.cfur * exp(-0.25 * rk2 / a.alp2) / rk2;
        const double s2 = sr * sr + si * si;
        const double akv = 0.5 * akc * (4.0 * a.alp2 + rk2) / (a.alp2 * rk2);
        e[0] = akc * s2;
        e[1] = akc * s2 - rxv * rxv * akv * s2;
        e[2] = akc * s2 - ryv * ryv * akv * s2;
        e[3] = akc * s2 - rzv * rzv * akv * s2;
        a.ak[kk] = make_double2(2.0 * akc * sr, 2.0 * akc * si);
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = e[q];
    __syncthreads();
    for (int s = kFurT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.epart[blockIdx.x] = make_double4(red[0][0], red[1][0], red[2][0], red[3][0]);
}

__global__ __launch_bounds__(kFurT) void k_fur_force(FurArgs a) {
    __shared__ int4 sk[kFurT];
    __shared__ double2 sa[kFurT];
    const int i = blockIdx.x * kFurT + threadIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0, qi = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
    if (i < a.n) {
        ux = a.x[i] / a.lx;
        uy = a.y[i] / a.ly;
        uz = a.z[i] / a.lz;
        qi = a.q[i];
    }
    for (int b = 0; b < a.nkv; b += kFurT) {
        const int kk = b + threadIdx.x;
        if (kk < a.nkv) {
            sk[threadIdx.x] = a.kv[kk];
            sa[threadIdx.x] = a.ak[kk];
        }
        __syncthreads();
        const int m = min(kFurT, a.nkv - b);
        for (int t = 0; t < m; ++t) {
            const int4 k = sk[t];
            double s, c;
            sincospi(2.0 * fur_turns(k, ux, uy, uz), &s, &c);
            // forces.f:1320-1326: QFOR = 2 AKC (Re S q sin - Im S q cos)
            const double qfor = qi * (sa[t].x * s - sa[t].y * c);
            fx += (double)k.x * qfor;
            fy += (double)k.y * qfor;
            fz += (double)k.z * qfor;
        }
        __syncthreads();
    }
    if (i < a.n) {
        const double twopi = 2.0 * kPI;
        a.fx[i] = twopi / a.lx * fx;
        a.fy[i] = twopi / a.ly * fy;
        a.fz[i] = twopi / a.lz * fz;
    }
}

__global__ __launch_bounds__(kFurT) void k_fur_esum(FurArgs a) {
    __shared__ double red[4][kFurT];
    const int nb = (a.nkv + kFurT - 1) / kFurT;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kFurT) {
        const double4 p = a.epart[b];
        e[0] += p.x;
        e[1] += p.y;
        e[2] += p.z;
        e[3] += p.w;
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = e[q];
    __syncthreads();
    for (int s = kFurT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (a.e4 && threadIdx.x < 4) a.e4[threadIdx.x] = red[threadIdx.x][0];
    if (a.qrow && threadIdx.x == 0) {
        for (int q = 0; q < Q_N; ++q) a.qrow[q] = 0.0;
        a.qrow[Q_PELS1] = red[0][0];
        a.qrow[Q_VIRX] = red[1][0];
        a.qrow[Q_VIRY] = red[2][0];
        a.qrow[Q_VIRZ] = red[3][0];
    }
}

// Workspace layout of one call (byte offsets, 256-aligned).
struct FurLayout {
    size_t kv, spart, ak, epart, count, total;
};

inline size_t fur_align(size_t b) { return (b + 255) / 256 * 256; }

inline FurLayout fur_layout(int n, int nkv) {
    const size_t nchunk = (size_t)(n + kFurChunk - 1) / kFurChunk;
    const size_t nkb = (size_t)(nkv + kFurT - 1) / kFurT;
    FurLayout L{};
    size_t o = 0;
    L.kv = o; o += fur_align(sizeof(int4) * (size_t)std::max(nkv, 1));
    L.spart = o; o += fur_align(sizeof(double2) * std::max<size_t>(1, nchunk * nkv));
    L.ak = o; o += fur_align(sizeof(double2) * (size_t)std::max(nkv, 1));
    L.epart = o; o += fur_align(sizeof(double4) * std::max<size_t>(nkb, 1));
    L.count = o; o += fur_align(sizeof(int));
    L.total = o;
    return L;
}

}  // namespace mdx
