// Neighbour-list construction (replaces CHCNB, forces.f:1813-2205).
//
// The reference scans all N^2/2 pairs.  Here atoms are binned into x-y
// columns, sorted by z inside each column and cut into clusters of 8; an
// i-cluster lists every j-cluster whose bounding box comes within the list
// radius (half list: every cluster pair once, cj >= ci).  For each listed cluster
// pair a wavefront tests all 64 atom pairs with the reference's arithmetic
// (FP64 differences of the caller's coordinates, one-shift minimum image,
// rr2 = (dx^2 + dy^2) + dz^2 without FMA contraction, inclusive <= against
// R_cut^2 and R_short^2), so the set of listed pairs and their short/long
// class are those CHCNB would produce.  Same-molecule pairs closer than 8 A
// are checked against the bound-atom list (forces.f:1943-1962): 1-2/1-3
// pairs are dropped, 1-4 pairs and intramolecular pairs of lambda-coupled
// molecules go to the special list (classified by k_special_classify).
#include "mdx_internal.h"

namespace mdx {

// Wrap into [0, L) and bin into x-y columns; key = column << 32 | z code.
// Each atom gets a rank in its column, unique (the column sort orders by
// (z code, atom index), so the ranks' order does not matter).  With
// per-block counts in LDS (lds_cols > 0, the column count) a block ranks
// its atoms with LDS atomics and adds each column's block count to the
// global count once; otherwise every atom takes a global atomic.
__global__ __launch_bounds__(256) void k_bin(int n, const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, Box b, int ncx, int ncy, int lds_cols,
                                             unsigned long long *__restrict__ keys, int *__restrict__ rank,
                                             int *__restrict__ colcount) {
    extern __shared__ int bcnt[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int c = threadIdx.x; c < lds_cols; c += blockDim.x) bcnt[c] = 0;
    int col = -1;
    if (i < n) {
        double w[3] = {x[i], y[i], z[i]};
        for (int k = 0; k < 3; ++k) {
            w[k] -= b.l[k] * floor(w[k] / b.l[k]);
            if (!(w[k] < b.l[k]) || w[k] < 0.0) w[k] = 0.0;
        }
        int cx = (int)(w[0] / b.l[0] * ncx), cy = (int)(w[1] / b.l[1] * ncy);
        cx = min(max(cx, 0), ncx - 1);
        cy = min(max(cy, 0), ncy - 1);
        col = cy * ncx + cx;
        double fz = w[2] / b.l[2] * 4294967295.0;
        unsigned int zq = (unsigned int)fmin(fmax(fz, 0.0), 4294967295.0);
        keys[i] = ((unsigned long long)col << 32) | zq;
    }
    if (lds_cols == 0) {
        if (col >= 0) rank[i] = atomicAdd(&colcount[col], 1);   // arrival rank in the column
        return;
    }
    __syncthreads();
    const int local = col >= 0 ? atomicAdd(&bcnt[col], 1) : 0;
    __syncthreads();
    for (int c = threadIdx.x; c < lds_cols; c += blockDim.x) {
        const int k = bcnt[c];
        if (k > 0) bcnt[c] = atomicAdd(&colcount[c], k);    // the block's base in the column
    }
    __syncthreads();
    if (col >= 0) rank[i] = bcnt[col] + local;
}

// Exclusive prefix sum of n ints, one block of 1,024 threads (each sums a
// contiguous chunk; chunk sums scanned per wavefront, then across the 16
// wavefronts); *total, when given, receives the sum of all n.
__global__ __launch_bounds__(1024) void k_scan_int(int n, const int *__restrict__ in, int *__restrict__ out,
                                                   int *__restrict__ total) {
    __shared__ int wtot[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    int sum = 0;
    for (int i = i0; i < i1; ++i) sum += in[i];
    int inc = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += wtot[k];
    int run = base + inc - sum;
    for (int i = i0; i < i1; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (total && t == 1023) *total = base + inc;
}

// Atoms into their column's range in arrival order, as the column-local
// sort key (z code << 32 | atom index): unique, and ordered as the stable
// sort by (column, z code) orders atoms of equal z code (by index).
__global__ void k_scatter(int n, const unsigned long long *__restrict__ keys, const int *__restrict__ rank,
                          const int *__restrict__ colstart, unsigned long long *__restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    const int col = (int)(k >> 32);
    tmp[colstart[col] + rank[i]] = ((k & 0xffffffffull) << 32) | (unsigned)i;
}

// One block per column: each atom's position in its column is the number
// of the column's keys below its own (keys staged through LDS 1,024 at a
// time); writes the sorted (column << 32 | z code) keys and atom indices.
__global__ __launch_bounds__(256) void k_col_sort(const int *__restrict__ colcount, const int *__restrict__ colstart,
                                                  const unsigned long long *__restrict__ tmp,
                                                  unsigned long long *__restrict__ skeys, int *__restrict__ svals) {
    __shared__ unsigned long long sk[1024];
    const int col = blockIdx.x, t = threadIdx.x;
    const int m = colcount[col], base = colstart[col];
    for (int g = 0; g < m; g += 256) {
        const int e = g + t;
        const unsigned long long ke = e < m ? tmp[base + e] : ~0ull;
        int cnt = 0;
        for (int c0 = 0; c0 < m; c0 += 1024) {
            const int c1 = min(m - c0, 1024);
            __syncthreads();
            for (int f = t; f < c1; f += 256) sk[f] = tmp[base + c0 + f];
            __syncthreads();
            for (int f = 0; f < c1; ++f) cnt += sk[f] < ke ? 1 : 0;
        }
        if (e < m) {
            skeys[base + cnt] = ((unsigned long long)col << 32) | (ke >> 32);
            svals[base + cnt] = (int)(ke & 0xffffffffull);
        }
    }
}

// Clusters per column: ceil(count / 8).
__global__ void k_col_clusters(int ncol, const int *__restrict__ colcount, int *__restrict__ colclu_n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol) colclu_n[c] = (colcount[c] + kClu - 1) / kClu;
}

// Sorted atom p -> cluster slot.
__global__ void k_fill(int n, const unsigned long long *__restrict__ skeys, const int *__restrict__ svals,
                       const int *__restrict__ colstart, const int *__restrict__ colclu,
                       int *__restrict__ clu_atom, int *__restrict__ atom_slot, int *__restrict__ clu_col) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int col = (int)(skeys[p] >> 32);
    const int r = p - colstart[col];
    const int c = colclu[col] + r / kClu;
    const int slot = c * kClu + (r % kClu);
    const int a = svals[p];
    clu_atom[slot] = a;
    atom_slot[a] = slot;
    if (r % kClu == 0) clu_col[c] = col;
}

// Cluster anchor (bounding-box centre of the wrapped coordinates, FP64)
// and half extents.
__global__ void k_cluster_geom(const int *__restrict__ pnclu, const int *__restrict__ clu_atom, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z, Box b,
                               double *__restrict__ anchor, float *__restrict__ ext, float *__restrict__ rbox) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *pnclu) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    double rlo[3] = {1e300, 1e300, 1e300}, rhi[3] = {-1e300, -1e300, -1e300};
    for (int s = 0; s < kClu; ++s) {
        const int a = clu_atom[c * kClu + s];
        if (a < 0) continue;
        double w[3] = {x[a], y[a], z[a]};
        for (int k = 0; k < 3; ++k) {
            rlo[k] = fmin(rlo[k], w[k]);
            rhi[k] = fmax(rhi[k], w[k]);
            w[k] -= b.l[k] * floor(w[k] / b.l[k]);
            lo[k] = fmin(lo[k], w[k]);
            hi[k] = fmax(hi[k], w[k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        anchor[c * 3 + k] = 0.5 * (lo[k] + hi[k]);
        ext[c * 3 + k] = (float)(0.5 * (hi[k] - lo[k])) * 1.0001f + 1e-4f;
        rbox[c * 6 + k] = __double2float_rd(rlo[k]);        // rounded outwards
        rbox[c * 6 + 3 + k] = __double2float_ru(rhi[k]);
    }
}

// Image code of a cluster pair (bits kCandCode + 2 q, axis q): the FP64
// differences x_i - x_j of the stored coordinates of all its atom pairs lie
// between the FP64 differences of the clusters' coordinate bounds (FP32,
// rounded outwards; rounding is monotonic), so when those fall on one side
// of the minimum-image test |d| > L / 2, the test has the same outcome for
// every pair: 0 none, 1 every d - L, 2 every d + L, 3 mixed (tested per
// pair).
__device__ inline unsigned image_code(const float (&rli)[3], const float (&rhi)[3], const float (&rj)[6], const Box &b) {
#pragma clang fp contract(off)
    unsigned code = 0;
    for (int q = 0; q < 3; ++q) {
        const double dmin = (double)rli[q] - (double)rj[3 + q], dmax = (double)rhi[q] - (double)rj[q];
        unsigned c = 3u;
        if (dmin > b.h[q]) c = 1u;
        else if (dmax < -b.h[q]) c = 2u;
        else if (dmax <= b.h[q] && dmin >= -b.h[q]) c = 0u;
        code |= c << (kCandCode + 2 * q);
    }
    return code;
}

// Candidate j-clusters of each i-cluster: bounding boxes within r_list,
// j >= i (a half list: every cluster pair once, the i-cluster's own pairs
// with i slot < j slot).  One wavefront per i-cluster; the clusters of each
// row of neighbour columns (periodic wrap, each column once) are tested 64
// at a time and the hits compacted in column order with a ballot, so the
// candidate order is that of a serial scan.
__global__ __launch_bounds__(64) void k_candidates(const int *__restrict__ pnclu, const double *__restrict__ anchor,
                                                   const float *__restrict__ ext, const float *__restrict__ rbox,
                                                   const int *__restrict__ clu_col,
                                                   const int *__restrict__ colclu, const int *__restrict__ colclu_n,
                                                   int ncx, int ncy, int rx, int ry, Box b, double rlist, int cap,
                                                   int *__restrict__ cand, int *__restrict__ ncand,
                                                   int *__restrict__ overflow) {
    const int ci = blockIdx.x;
    if (ci >= *pnclu) return;
    const int lane = threadIdx.x;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int col = clu_col[ci];
    const int cx0 = col % ncx, cy0 = col / ncx;
    const double ai[3] = {anchor[ci * 3], anchor[ci * 3 + 1], anchor[ci * 3 + 2]};
    const float ei[3] = {ext[ci * 3], ext[ci * 3 + 1], ext[ci * 3 + 2]};
    const double il[3] = {1.0 / b.l[0], 1.0 / b.l[1], 1.0 / b.l[2]};
    const int nx = min(2 * rx + 1, ncx), ny = min(2 * ry + 1, ncy);
    const int sx = (2 * rx + 1 >= ncx) ? 0 : cx0 - rx + ncx;
    const int sy = (2 * ry + 1 >= ncy) ? 0 : cy0 - ry + ncy;
    const double r2 = rlist * rlist;
    const float rli[3] = {rbox[ci * 6], rbox[ci * 6 + 1], rbox[ci * 6 + 2]};
    const float rhi[3] = {rbox[ci * 6 + 3], rbox[ci * 6 + 4], rbox[ci * 6 + 5]};
    int cnt = 0;
    // clusters [lo, hi) tested 64 at a time, hits compacted in index order
    // (j >= i: clusters below ci, which include every cluster of a column
    // numbered below the i-cluster's, are left out)
    auto scan = [&](int lo, int hi) {
        for (int base = max(lo, ci); base < hi; base += 64) {
            const int cj = base + lane;
            bool hit = false;
            float rj[6];                    // loaded with the anchor (one round of loads)
            if (cj < hi) {
                double aj[3];
                float ej[3];
                for (int k = 0; k < 3; ++k) {
                    aj[k] = anchor[cj * 3 + k];
                    ej[k] = ext[cj * 3 + k];
                }
                for (int k = 0; k < 6; ++k) rj[k] = rbox[cj * 6 + k];
                double d2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    double d = ai[k] - aj[k];
                    d -= b.l[k] * rint(d * il[k]);
                    const double g = fabs(d) - (double)(ei[k] + ej[k]);
                    if (g > 0.0) d2 += g * g;
                }
                hit = d2 <= r2;
            }
            const unsigned long long m = __ballot(hit);
            const int pos = cnt + __popcll(m & below);
            if (hit && pos < cap) cand[(size_t)ci * cap + pos] = (int)((unsigned)cj | image_code(rli, rhi, rj, b));
            cnt += __popcll(m);
        }
    };
    // clusters are numbered column by column, so the neighbour columns of
    // one row are one index range, or two where they wrap around x (first
    // the columns from the left edge of the neighbourhood to the box's
    // edge, then those from x = 0): the order of a serial scan column by
    // column
    auto cols = [&](int c0, int c1) { return make_int2(colclu[c0], colclu[c1] + colclu_n[c1]); };
    for (int oy = 0; oy < ny; ++oy) {
        const int cy = (sy + oy) % ncy;
        const int r0 = cy * ncx;
        if (r0 + ncx - 1 < col) continue;       // every column of the row below the i-cluster's
        const int x0 = sx % ncx;
        if (x0 + nx <= ncx) {
            const int2 r = cols(r0 + x0, r0 + x0 + nx - 1);
            scan(r.x, r.y);
        } else {
            const int2 ra = cols(r0 + x0, r0 + ncx - 1), rb = cols(r0, r0 + x0 + nx - ncx - 1);
            scan(ra.x, ra.y);
            scan(rb.x, rb.y);
        }
    }
    if (lane == 0) {
        ncand[ci] = min(cnt, cap);
        if (cnt > cap) atomicMax(overflow, cnt);
    }
}

// Is j bound to i (returns 1 for 1-2/1-3, 2 for 1-4, 0 otherwise)?
__device__ inline int bound_kind(int i, int j, const int *__restrict__ excl_start,
                                 const int *__restrict__ excl_atom, const signed char *__restrict__ excl_is14) {
    for (int k = excl_start[i]; k < excl_start[i + 1]; ++k)
        if (excl_atom[k] == j) return excl_is14[k] ? 2 : 1;
    return 0;
}

// Pair masks for every candidate of an i-cluster; one wavefront per
// i-cluster, lane = i_slot * 8 + j_slot.
__global__ __launch_bounds__(64) void k_masks(
    const int *__restrict__ pnclu, const int *__restrict__ cand, const int *__restrict__ ncand, int cap,
    const int *__restrict__ clu_atom, const double4 *__restrict__ spos, const double *__restrict__ anchor,
    const float *__restrict__ ext, double rlist, const int *__restrict__ atom_moltype, const int *__restrict__ type_ee, const int *__restrict__ excl_start,
    const int *__restrict__ excl_atom, const signed char *__restrict__ excl_is14, Box b, double rl2,
    double short2, PairEntry *__restrict__ list, int *__restrict__ nlist, int *__restrict__ nshort,
    int list_cap, int *__restrict__ overflow) {
#pragma clang fp contract(off)
    const int ci = blockIdx.x;
    if (ci >= *pnclu) return;
    const int lane = threadIdx.x;
    const int il = lane >> 3, jl = lane & 7;
    const int ai = clu_atom[ci * kClu + il];
    const double4 pi = spos[ci * kClu + il];
    const int moli = (int)pi.w;
    const int eei = ai >= 0 ? type_ee[atom_moltype[ai]] : 1;
    const double aci[3] = {anchor[ci * 3], anchor[ci * 3 + 1], anchor[ci * 3 + 2]};
    const double eci[3] = {ext[ci * 3], ext[ci * 3 + 1], ext[ci * 3 + 2]};
    const double ilb[3] = {1.0 / b.l[0], 1.0 / b.l[1], 1.0 / b.l[2]};
    const int nc = ncand[ci];
    const MDX_CONST int *cl = sld(cand) + (size_t)ci * cap;
    // software pipeline: candidate indices two ahead (scalar loads), the j
    // slot's coordinates one ahead; the loads are unconditional (past the
    // end they re-read the last candidate) so none is waited for at once
    const int nl = max(nc - 1, 0);
    int c1 = nc > 0 ? cl[0] : ci, c2 = cl[min(1, nl)];
    if (nc == 0) c2 = ci;
    double4 p1 = spos[(int)((unsigned)c1 & kCandIdx) * kClu + jl];
    // entries with a short-class pair fill the list from the front, the
    // others from the back (in discovery order both), so a fast-class call
    // reads only the front segment.  A listed candidate's (j-cluster,
    // masks, list slot) go into lane nb of a buffer (a select per value); every 64
    // of them, each lane builds its entry (image shift, per-pair-image
    // flag) with its own loads of the j-cluster's anchor and extents.
    int cnt = 0, cnt_s = 0;
    int nb = 0;
    int bcj = 0, bslot = 0, bm0 = 0, bm1 = 0, bs0 = 0, bs1 = 0;
    auto emit = [&]() {
        if (lane < nb) {
            PairEntry e;
            double s[3];
            bool wrap = false;
            for (int q = 0; q < 3; ++q) {
                // d * (1/L) and d / L round to the same image unless the
                // anchors are near half a box apart; such entries take the
                // minimum image per pair, so either image serves
                const double d = aci[q] - anchor[bcj * 3 + q];
                s[q] = d - b.l[q] * rint(d * ilb[q]);
                wrap = wrap || rlist + eci[q] + (double)ext[bcj * 3 + q] + kImageMargin >= b.h[q];
            }
            e.cj = (int)((unsigned)bcj | (wrap ? kEntryWrap : 0u));
            e.sx = (float)s[0]; e.sy = (float)s[1]; e.sz = (float)s[2];
            e.mask = ((unsigned long long)(unsigned)bm1 << 32) | (unsigned)bm0;
            e.mshort = ((unsigned long long)(unsigned)bs1 << 32) | (unsigned)bs0;
            list[(size_t)ci * list_cap + bslot] = e;
        }
        nb = 0;
    };
    for (int k = 0; k < nc; ++k) {
        const int cj = (int)((unsigned)c1 & kCandIdx);
        const unsigned code = (unsigned)c1 >> kCandCode;   // image code (image_code)
        const double4 pj = p1;
        c1 = c2;
        p1 = spos[(int)((unsigned)c1 & kCandIdx) * kClu + jl];
        c2 = cl[min(k + 2, nl)];
        bool listed = false, shrt = false;
        // an empty slot has molecule -1; (cj != ci || il < jl) keeps each
        // pair once and leaves out an atom with itself
        if (moli >= 0 && pj.w >= 0.0 && (cj != ci || il < jl)) {
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            // the reference's minimum image: per pair where the cluster
            // pair's code leaves it open, otherwise its known outcome (the
            // same subtraction, so the same rounding)
            if (code != 0u) {
                if ((code & (code >> 1) & 0x15u) != 0u) min_image(dx, dy, dz, b);
                else {
                    const unsigned cx = code & 3u, cy = (code >> 2) & 3u, cz = (code >> 4) & 3u;
                    if (cx) dx = cx == 1u ? dx - b.l[0] : dx + b.l[0];
                    if (cy) dy = cy == 1u ? dy - b.l[1] : dy + b.l[1];
                    if (cz) dz = cz == 1u ? dz - b.l[2] : dz + b.l[2];
                }
            }
            double rr2 = dx * dx + dy * dy;
            rr2 = rr2 + dz * dz;
            listed = rr2 <= rl2;
            shrt = rr2 <= short2;
            if (listed && moli == (int)pj.w) {
                const int aj = clu_atom[cj * kClu + jl];
                if (rr2 <= 64.0 && bound_kind(ai, aj, excl_start, excl_atom, excl_is14) != 0) listed = false;
                else if (eei == 0) listed = false;
            }
        }
        const unsigned long long m = __ballot(listed);
        const unsigned long long ms = __ballot(listed && shrt);
        if (m != 0ull) {
            const int slot = ms != 0ull ? cnt_s : list_cap - 1 - (cnt - cnt_s);
            if (cnt < list_cap) {
                const bool mine = lane == nb;
                bcj = mine ? cj : bcj;
                bslot = mine ? slot : bslot;
                bm0 = mine ? (int)(unsigned)m : bm0;
                bm1 = mine ? (int)(unsigned)(m >> 32) : bm1;
                bs0 = mine ? (int)(unsigned)ms : bs0;
                bs1 = mine ? (int)(unsigned)(ms >> 32) : bs1;
                if (++nb == 64) emit();
            }
            ++cnt;
            if (ms != 0ull) ++cnt_s;
        }
    }
    emit();
    if (lane == 0) {
        nlist[ci] = min(cnt, list_cap);
        nshort[ci] = min(cnt_s, list_cap);
        if (cnt > list_cap) atomicMax(overflow, cnt);
    }
}

// Transposed list: the entries that name cluster c as their j-cluster get
// consecutive records from jstart[c] (in units of 24 floats), so the pair
// kernel writes each entry's j-reaction record where k_jgather streams it
// for the j atoms, without atomics.  A segment holds the records of the
// front-segment entries first (both classes, 2 units each, jcnt.x of them),
// then those of the long-class entries (1 unit each, jcnt.y).  k_jcount
// gives each entry its rank in its part (integer atomics, at rebuilds
// only), k_scan_excl turns the unit counts into segment starts, and
// k_jplace turns each entry's rank into its record position (jdest, per
// list slot).  The order inside each part follows the atomics; the sums
// over it are integer sums, so results do not depend on that order.
__global__ __launch_bounds__(64) void k_jcount(const int *__restrict__ pnclu, const PairEntry *__restrict__ list,
                                               const int *__restrict__ nlist, const int *__restrict__ nshort,
                                               int list_cap, int2 *__restrict__ jcnt, unsigned *__restrict__ jrank) {
    const int ci = blockIdx.x;
    if (ci >= *pnclu) return;
    const int n = nlist[ci], ns = nshort[ci];
    for (int k = threadIdx.x; k < n; k += 64) {
        const size_t p = (size_t)ci * list_cap + (k < ns ? k : list_cap - 1 - (k - ns));
        const int cj = (int)((unsigned)list[p].cj & ~kEntryWrap);
        jrank[p] = (unsigned)atomicAdd(k < ns ? &jcnt[cj].x : &jcnt[cj].y, 1);
    }
}

__global__ __launch_bounds__(64) void k_jplace(const int *__restrict__ pnclu, const PairEntry *__restrict__ list,
                                               const int *__restrict__ nlist, const int *__restrict__ nshort,
                                               int list_cap, const unsigned *__restrict__ jstart,
                                               const int2 *__restrict__ jcnt, unsigned *__restrict__ jdest) {
    const int ci = blockIdx.x;
    if (ci >= *pnclu) return;
    const int n = nlist[ci], ns = nshort[ci];
    for (int k = threadIdx.x; k < n; k += 64) {
        const size_t p = (size_t)ci * list_cap + (k < ns ? k : list_cap - 1 - (k - ns));
        const int cj = (int)((unsigned)list[p].cj & ~kEntryWrap);
        const unsigned r = jdest[p];
        jdest[p] = jstart[cj] + (k < ns ? 2u * r : 2u * (unsigned)jcnt[cj].x + r);
    }
}

// Exclusive prefix sum of the record units 2 x + y of n clusters, one block
// of 1,024 threads: thread t
// sums a contiguous chunk, the chunk sums are scanned across the block
// (wavefront scans, then the 16 wavefront totals), and each thread writes
// its chunk's prefixes.
__global__ __launch_bounds__(1024) void k_scan_excl(const int *__restrict__ pn, const int2 *__restrict__ in,
                                                    unsigned *__restrict__ out) {
    __shared__ unsigned wtot[16];
    const int n = *pn;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    unsigned sum = 0;
    for (int i = i0; i < i1; ++i) sum += 2u * (unsigned)in[i].x + (unsigned)in[i].y;
    unsigned inc = sum;   // inclusive scan inside the wavefront
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    unsigned base = 0;
    for (int k = 0; k < w; ++k) base += wtot[k];
    unsigned run = base + inc - sum;
    for (int i = i0; i < i1; ++i) {
        const unsigned v = 2u * (unsigned)in[i].x + (unsigned)in[i].y;
        out[i] = run;
        run += v;
    }
}

// Per-slot FP64 coordinates (the caller's, as the reference differences
// them) and molecule index (-1: empty slot) for the mask pass.
__global__ void k_slot_pos(const int *__restrict__ pnclu, const int *__restrict__ clu_atom, const double *__restrict__ x,
                           const double *__restrict__ y, const double *__restrict__ z,
                           const int *__restrict__ atom_mol, double4 *__restrict__ spos) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= *pnclu * kClu) return;
    const int a = clu_atom[s];
    spos[s] = a >= 0 ? make_double4(x[a], y[a], z[a], (double)atom_mol[a]) : make_double4(0.0, 0.0, 0.0, -1.0);
}

// Special pairs (1-4 pairs, intramolecular pairs of lambda-coupled
// molecules) re-classified from the current coordinates with the same
// arithmetic and rules as CHCNB.  Candidates are (i > j) pairs fixed by the
// topology; kind bit 0: bound 1-2/1-3, bit 1: bound 1-4, bit 2: EE molecule.
__global__ void k_special_classify(int ns, const int2 *__restrict__ pairs, const signed char *__restrict__ kind,
                                   const double *__restrict__ x, const double *__restrict__ y,
                                   const double *__restrict__ z, Box b, double rl2, double short2,
                                   signed char *__restrict__ cls) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int i = pairs[k].x, j = pairs[k].y;
    double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
    min_image(dx, dy, dz, b);
    double rr2 = dx * dx + dy * dy;
    rr2 = rr2 + dz * dz;
    const int kd = kind[k];
    signed char c = SP_NONE;
    if (rr2 <= 64.0 && (kd & 1)) c = SP_NONE;
    else if (rr2 <= 64.0 && (kd & 2)) c = SP_SHORT14;
    else if (kd & 4) c = rr2 <= short2 ? SP_SHORT : (rr2 <= rl2 ? SP_LONG : SP_NONE);
    cls[k] = c;
}

}  // namespace mdx
