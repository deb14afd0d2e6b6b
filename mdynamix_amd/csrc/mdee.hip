// Expanded-ensemble energies at trial coupling parameters (the energy
// differences of a lambda move: DIFEN mdee.f:379-530, DIFFUR mdee.f:265-370),
// in FP64, for a list of lambda values at once.  Each sum is a function of
// lambda alone (the reference's "old" energies use the parameters RECINT set
// from the current lambda, mdee.f:86-137, which are these formulas at it).
//
// k_ee_pairs   one wavefront per i-cluster of the cluster-pair list, lane
//              (i_slot * 8 + j_slot): the listed pairs with a lambda-coupled
//              atom in different molecules (entries without one are skipped
//              wave-uniformly); per pair and lambda the soft-core Coulomb
//              erfc(alpha r) / (r + shift) and LJ terms; per-lane FP64 sums
//              in LDS, one row of nlam sums per i-cluster
// k_ee_colsum  one block per lambda: the rows (or the k-block partials) in
//              order
// k_ee_sk      FURIR-like structure factors in FP64 with two charge sets:
//              the uncoupled atoms' current charges and the coupled atoms'
//              unscaled charges (S(lambda) = S_a + lambda^PW1 S_b)
// k_ee_sred    per k: AKC |S_a + s S_b|^2 for every lambda, block partials
// k_ee_intra   one block: the pairs inside lambda-coupled molecules,
//              COULF q0_i q0_j (erfc(alpha r) - 1) / r (DIFFUR's
//              intramolecular term at lambda = 1)
#include "mdx_internal.h"

namespace mdx {

constexpr int kEeMaxLam = 32;

struct EeArgs {
    int nclu, list_cap, nlam;
    const PairEntry *list;
    const int *nlist, *nshort, *clu_atom;
    const double *x, *y, *z;
    Box box;
    double rc2d;                      // cutoff test (Verlet lists) or +inf
    const int *atom_mol, *atom_moltype, *atom_ljtype, *type_ee;
    const double *q0;                 // CHARGE per atom
    const double *a6_0, *b12_0;       // nlj x nlj, unscaled
    int nlj;
    double coulf, alpha;
    const double *lam;                // [2 nlam]: lambda^PW1, BEE (1 - lambda^PW2)
    double *rows;                     // [nclu][nlam]
};

__device__ inline double ee_erfc(double ar) {
    const double t = 1.0 / (1.0 + kB1 * ar);
    return ((((kA5 * t + kA4) * t + kA3) * t + kA2) * t + kA1) * t * exp(-(ar * ar));
}

__global__ __launch_bounds__(64) void k_ee_pairs(EeArgs a) {
#pragma clang fp contract(off)
    __shared__ double s_acc[kEeMaxLam * 64];
    const int ci = blockIdx.x, lane = threadIdx.x;
    const int il = lane >> 3, jl = lane & 7;
    for (int L = 0; L < a.nlam; ++L) s_acc[L * 64 + lane] = 0.0;
    const int ai = a.clu_atom[ci * kClu + il];
    const bool eei = ai >= 0 && a.type_ee[a.atom_moltype[ai]] == 0;
    const int moli = ai >= 0 ? a.atom_mol[ai] : -1;
    const int lji = ai >= 0 ? a.atom_ljtype[ai] : 0;
    const double xi = ai >= 0 ? a.x[ai] : 0.0, yi = ai >= 0 ? a.y[ai] : 0.0, zi = ai >= 0 ? a.z[ai] : 0.0;
    const int ns = a.nshort[ci], n = a.nlist[ci];
    const PairEntry *Lst = a.list + (size_t)ci * a.list_cap;
    for (int k = 0; k < n; ++k) {
        const PairEntry e = Lst[k < ns ? k : a.list_cap - 1 - (k - ns)];
        const int cj = (int)((unsigned)e.cj & ~kEntryWrap);
        const bool listed = (e.mask >> lane) & 1ull;
        const int aj = listed ? a.clu_atom[cj * kClu + jl] : -1;
        const bool eej = aj >= 0 && a.type_ee[a.atom_moltype[aj]] == 0;
        const bool take = listed && (eei || eej) && a.atom_mol[aj] != moli;
        if (!__any(take)) continue;
        if (!take) continue;
        double dx = xi - a.x[aj], dy = yi - a.y[aj], dz = zi - a.z[aj];
        min_image(dx, dy, dz, a.box);
        double rr = dx * dx + dy * dy;
        rr = rr + dz * dz;
        if (!(rr <= a.rc2d)) continue;
        const double r = sqrt(rr);
        const double erfcr = ee_erfc(a.alpha * r);
        const double c = a.q0[ai] * a.q0[aj] * a.coulf;
        const int t = lji * a.nlj + a.atom_ljtype[aj];
        const double a6 = a.a6_0[t], b12 = a.b12_0[t];
        const bool both = eei && eej;
        for (int L = 0; L < a.nlam; ++L) {
            const double s = a.lam[2 * L], sh = a.lam[2 * L + 1];
            const double qs = both ? s * s : s;
            const double r1ni = 1.0 / (r + sh);
            const double ees = qs * c * erfcr * r1ni;
            const double r2 = r1ni * r1ni;
            const double r6 = r2 * r2 * r2;
            const double esr = (s * a6 + s * b12 * r6) * r6;
            s_acc[L * 64 + lane] += ees + esr;
        }
    }
    for (int L = 0; L < a.nlam; ++L) {
        double v = s_acc[L * 64 + lane];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) a.rows[(size_t)ci * a.nlam + L] = v;
    }
}

// Sum of rows[r * stride + col] over r < nrows in a fixed order (block of
// 256 threads; column = blockIdx.x).
__global__ __launch_bounds__(256) void k_ee_colsum(int nrows, int stride, const double *rows, double *out) {
    __shared__ double red[256];
    const int col = blockIdx.x;
    double v = 0.0;
    for (int r = threadIdx.x; r < nrows; r += 256) v += rows[(size_t)r * stride + col];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = red[0];
}

struct EeFurArgs {
    int n, nkv, nchunk, nlam;
    double lx, ly, lz, alp2, cfur;
    const double *x, *y, *z;
    const double *q, *q0;             // current charges, CHARGE
    const int *atom_moltype, *type_ee;
    const int4 *kv;                   // [nkv] (n_x, n_y, n_z)
    double4 *spart;                   // [nchunk][nkv]: S_a (re, im), S_b (re, im)
    const double *lam;                // [2 nlam]
    double *kpart;                    // [nkb][nlam]
};

__global__ __launch_bounds__(kFurT) void k_ee_sk(EeFurArgs a) {
    __shared__ double sx[kFurT], sy[kFurT], sz[kFurT], sqa[kFurT], sqb[kFurT];
    const int kk = blockIdx.x * kFurT + threadIdx.x;
    const int ch = blockIdx.y;
    const int a0 = ch * kFurChunk, a1 = min(a0 + kFurChunk, a.n);
    const int4 k = kk < a.nkv ? a.kv[kk] : make_int4(0, 0, 0, 0);
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    for (int b = a0; b < a1; b += kFurT) {
        const int i = b + threadIdx.x;
        if (i < a1) {
            sx[threadIdx.x] = a.x[i] / a.lx;
            sy[threadIdx.x] = a.y[i] / a.ly;
            sz[threadIdx.x] = a.z[i] / a.lz;
            const bool ee = a.type_ee[a.atom_moltype[i]] == 0;
            sqa[threadIdx.x] = ee ? 0.0 : a.q[i];
            sqb[threadIdx.x] = ee ? a.q0[i] : 0.0;
        }
        __syncthreads();
        const int m = min(kFurT, a1 - b);
        for (int t = 0; t < m; ++t) {
            double s, c;
            sincospi(2.0 * fur_turns(k, sx[t], sy[t], sz[t]), &s, &c);
            ar += sqa[t] * c;
            ai += sqa[t] * s;
            br += sqb[t] * c;
            bi += sqb[t] * s;
        }
        __syncthreads();
    }
    if (kk < a.nkv) a.spart[(size_t)ch * a.nkv + kk] = make_double4(ar, ai, br, bi);
}

__global__ __launch_bounds__(kFurT) void k_ee_sred(EeFurArgs a) {
    __shared__ double red[kFurT];
    const int kk = blockIdx.x * kFurT + threadIdx.x;
    double sar = 0.0, sai = 0.0, sbr = 0.0, sbi = 0.0, akc = 0.0;
    if (kk < a.nkv) {
        for (int c = 0; c < a.nchunk; ++c) {
            const double4 p = a.spart[(size_t)c * a.nkv + kk];
            sar += p.x; sai += p.y; sbr += p.z; sbi += p.w;
        }
        const int4 k = a.kv[kk];
        const double twopi = 2.0 * kPI;
        const double rxv = twopi * k.x / a.lx, ryv = twopi * k.y / a.ly, rzv = twopi * k.z / a.lz;
        double rk2 = rxv * rxv + ryv * ryv;
        rk2 = rk2 + rzv * rzv;
        akc = a.cfur * exp(-0.25 * rk2 / a.alp2) / rk2;
    }
    for (int L = 0; L < a.nlam; ++L) {
        const double s = a.lam[2 * L];
        const double cr = sar + s * sbr, cim = sai + s * sbi;
        red[threadIdx.x] = kk < a.nkv ? (cim * cim + cr * cr) * akc : 0.0;
        __syncthreads();
        for (int w = kFurT / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.kpart[(size_t)blockIdx.x * a.nlam + L] = red[0];
        __syncthreads();
    }
}

// Pairs (i, j) inside lambda-coupled molecules: sum of COULF q0_i q0_j
// (erfc(alpha r) - 1) / r (one-shift minimum image), fixed order.
__global__ __launch_bounds__(256) void k_ee_intra(int np, const int2 *pairs, const double *x, const double *y,
                                                  const double *z, const double *q0, Box box, double coulf,
                                                  double alpha, double *out) {
    __shared__ double red[256];
    double v = 0.0;
    for (int p = threadIdx.x; p < np; p += 256) {
#pragma clang fp contract(off)
        const int i = pairs[p].x, j = pairs[p].y;
        double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
        min_image(dx, dy, dz, box);
        double r2 = dx * dx + dy * dy;
        r2 = r2 + dz * dz;
        const double r1 = sqrt(r2);
        v += coulf * q0[i] * q0[j] * (ee_erfc(alpha * r1) - 1.0) * (r1 / r2);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace mdx
