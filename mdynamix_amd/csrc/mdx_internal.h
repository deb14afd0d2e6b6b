// Internal declarations of the MI355X force engine (gfx950, wave64).
//
// Data layout in HBM (one engine handle = one system):
//   atoms      FP64 SoA x/y/z as given by the caller (reference coordinates)
//   clusters   8 atoms each, built at every list rebuild from x-y columns
//              sorted by z; per cluster an FP64 anchor and per slot a packed
//              float4 (x-anchor, y-anchor, z-anchor, q) plus an int info word
//              (Lennard-Jones type | ee flag << 16); empty slots hold q = 0 and
//              never appear in a pair mask
//   pair list  full cluster-pair list (each i-cluster owns its j-clusters),
//              32 bytes per entry: j-cluster, FP32 image shift, 64-bit mask
//              of listed pairs and 64-bit mask of short-class pairs (bit
//              i_slot * 8 + j_slot)
//   specials   atom pairs the cluster kernel does not handle: 1-4 pairs and
//              intramolecular pairs of lambda-coupled molecules, classified
//              short/long at every rebuild
//   forces     pair forces FP64 per atom (plain stores, one owner per atom);
//              scattered terms (bonded, specials, exclusion corrections)
//              in 64-bit fixed point with integer atomics, so every result is
//              independent of execution order
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdx {

// prcm.h:27-32 (Abramowitz-Stegun 7.1.26 erfc and the reference's pi)
constexpr double kA1 = 0.254829592, kA2 = -0.284496736, kA3 = 1.421413741;
constexpr double kA4 = -1.453152027, kA5 = 1.061405429, kB1 = 0.3275911;
constexpr double kPI = 3.1415926536, kTOTPI = 1.128379169;

constexpr int kClu = 8;            // atoms per cluster
constexpr int kWave = 64;          // wavefront size on gfx950

// Energy/virial slots of the per-block partial sums.
enum EQ {
    Q_PE1 = 0, Q_PE2, Q_PELS1, Q_PELS2, Q_SPE, Q_PINT, Q_VIR1, Q_VIR2, Q_VIRB,
    Q_VIRX, Q_VIRY, Q_VIRZ, Q_VIRFX, Q_VIRFY, Q_VIRFZ, Q_WIRS, Q_WIRSS, Q_NPS, Q_NPL, Q_N
};

struct PairEntry {          // 32 bytes, read with scalar loads
    int32_t cj;
    float sx, sy, sz;
    uint64_t mask;          // pairs in the list
    uint64_t mshort;        // of those, the short (fast) class
};

// Special pair classes.
enum { SP_NONE = 0, SP_SHORT14 = 1, SP_SHORT = 2, SP_LONG = 3 };

struct Box {
    double l[3], h[3];      // edges and half edges
};

__host__ __device__ inline double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Single-shift minimum image of the reference pair loops (forces.f:1841-1867).
__device__ inline void min_image(double &dx, double &dy, double &dz, const Box &b) {
    if (fabs(dx) > b.h[0]) dx = dx - fsign(b.l[0], dx);
    if (fabs(dz) > b.h[2]) dz = dz - fsign(b.l[2], dz);
    if (fabs(dy) > b.h[1]) dy = dy - fsign(b.l[1], dy);
}

// Fixed-point force accumulation.
__device__ inline void fx_add(unsigned long long *p, double v, double scale) {
    long long q = __double2ll_rn(v * scale);
    atomicAdd(p, (unsigned long long)q);
}

}  // namespace mdx
