// Internal declarations of the MI355X force engine (gfx950, wave64).
//
// Data layout in HBM (one engine handle = one system):
//   atoms      FP64 SoA x/y/z as given by the caller (reference coordinates)
//   clusters   8 atoms each, built at every list rebuild from x-y columns
//              sorted by z; per cluster an FP64 anchor and per slot a packed
//              float4 (x-anchor, y-anchor, z-anchor, q) plus an int2 (LJ type
//              | ee flag << 16, low part of q); empty slots hold q = 0 and
//              never appear in a pair mask
//   pair list  half cluster-pair list (each cluster pair once, under the
//              lower-numbered i-cluster), 32 bytes per entry: j-cluster (and
//              the per-pair image flag), FP32 image shift, 64-bit mask of
//              listed pairs and 64-bit mask of short-class pairs (bit
//              i_slot * 8 + j_slot)
//   specials   atom pairs the cluster kernel does not handle: 1-4 pairs and
//              intramolecular pairs of lambda-coupled molecules, classified
//              short/long at every rebuild
//   forces     pair forces FP64 per atom (plain stores, one owner per atom);
//              scattered terms (bonded, specials, exclusion corrections,
//              bead springs) in 64-bit fixed point with integer atomics, so
//              every result is independent of execution order
//
// Every kernel is written as a device body taking its argument struct and a
// block index; the __global__ entry points read the struct of their system
// from an array in HBM through a block map (system, local block), so one
// launch covers every system of a batch (PIMD beads, expanded-ensemble
// jobs) and a single system is a batch of one.  Counts that change at a
// list rebuild live in the argument structs, launch shapes use capacities,
// and a captured hipGraph stays valid across rebuilds.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdx {

// prcm.h:27-32 (Abramowitz-Stegun 7.1.26 erfc and the reference's pi)
constexpr double kA1 = 0.254829592, kA2 = -0.284496736, kA3 = 1.421413741;
constexpr double kA4 = -1.453152027, kA5 = 1.061405429, kB1 = 0.3275911;
constexpr double kPI = 3.1415926536, kTOTPI = 1.128379169;

constexpr int kClu = 8;            // atoms per cluster
constexpr int kRecUnit = 24;       // floats per unit of the j-reaction records
// Per-type energy breakdown (MDX_WANT_DETAIL): per partial row
// [POTE1(ntp) | POTL1(ntp) | PES141(nmt) | PSR141(nmt)]
constexpr int kMaxDet = 2 * 137 + 2 * 16;
constexpr int kWave = 64;          // wavefront size on gfx950
constexpr int kForceTiles = 3;     // 16-atom tiles per wavefront of the reciprocal force pass
constexpr int kForceBlock = 64 * kForceTiles;   // atoms per block (4 wavefronts) of that pass
constexpr int kComAtoms = 512;     // LDS stage (atoms) of the centre-of-mass kernels
constexpr int kSkRows = 64;        // k rows per structure-factor block (16 per wavefront)
constexpr int kSkSub = 32;         // atoms per LDS stage of the structure-factor block
constexpr int kSkRedItems = 64;    // pair entries / rows per block of k_sk_reduce (four lanes each)
constexpr int kRedRows = 192;      // energy rows per block of the final reduction (16 x 256 / Q_N)

// Energy/virial slots of the per-block partial sums.
enum EQ {
    Q_PE1 = 0, Q_PE2, Q_PELS1, Q_PELS2, Q_SPE, Q_PINT, Q_VIR1, Q_VIR2, Q_VIRB,
    Q_VIRX, Q_VIRY, Q_VIRZ, Q_VIRFX, Q_VIRFY, Q_VIRFZ, Q_WIRS, Q_WIRSS, Q_NPS, Q_NPL,
    Q_QEK, Q_VIRPI, Q_N
};

struct PairEntry {          // 32 bytes, read with scalar loads
    int32_t cj;             // j-cluster; bit 31 (kEntryWrap): minimum image per pair
    float sx, sy, sz;
    uint64_t mask;          // pairs in the list
    uint64_t mshort;        // of those, the short (fast) class
};

// A list entry takes the minimum image per pair (bit 31 of PairEntry.cj)
// unless, on every axis, r_list + the two clusters' half extents +
// kImageMargin stays below half the box: then every atom pair within reach
// has the image of the two cluster anchors, also after each atom has moved
// up to kImageMargin / 4 since the build (the bound adds twice the motion
// to the pair distance and twice to the extents).
constexpr uint32_t kEntryWrap = 0x80000000u;
constexpr double kImageMargin = 4.0;
// Candidate list word: j-cluster index in the low kCandCode bits, the pair's
// image code (2 bits per axis) above.
constexpr int kCandCode = 26;
constexpr unsigned kCandIdx = (1u << kCandCode) - 1u;

// Special pair classes.
enum { SP_NONE = 0, SP_SHORT14 = 1, SP_SHORT = 2, SP_LONG = 3 };

struct Box {
    double l[3], h[3];      // edges and half edges
};

struct RecipRow {           // 16 bytes
    int nx, ny, nz0, len;
};

// Rows of the force pass, in (n_x, n_y) order so consecutive rows of one
// n_x differ by one in n_y.  Entries j = 1..m pair the k vectors with
// n_z = +j and -j (absent ones count as A = 0); k0 is the n_z = 0 vector
// or -1.
struct FRow {               // 32 bytes
    int ix, iy, m, off, k0, pad0, pad1, pad2;
};

// Force-pass tile: up to 4 rows of one n_x with consecutive n_y; its matrix
// operand (nstep steps of 16 values of j, kForceStepH halves each) starts at
// half woff and is scaled by 2^pad0.
struct FTile {              // 32 bytes
    int ix, iy0, nrow, nstep, woff, r0, pad0, pad1;
};
constexpr int kForceStepH = 1024;  // halves per force-pass step: 64 lanes x 8 (hi), then 64 x 8 (lo)

__host__ __device__ inline double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Address-space views of the pointers kept in argument structs (which the
// kernels read from HBM, so the compiler cannot infer where they point):
// gld/gst for per-lane global accesses, sld for wavefront-uniform
// read-only data, which then goes through scalar loads.
#define MDX_GLOBAL __attribute__((address_space(1)))
#define MDX_CONST __attribute__((address_space(4)))
template <typename T>
__device__ inline const MDX_GLOBAL T *gld(const T *p) { return (const MDX_GLOBAL T *)p; }
template <typename T>
__device__ inline MDX_GLOBAL T *gst(T *p) { return (MDX_GLOBAL T *)p; }
template <typename T>
__device__ inline const MDX_CONST T *sld(const T *p) { return (const MDX_CONST T *)p; }
// Value of a struct or vector type through such a view, word by word (C++
// copy constructors bind generic references only).
template <typename T>
__device__ inline T ldv(const MDX_CONST T *p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized types only");
    T v;
    int *d = reinterpret_cast<int *>(&v);
    const MDX_CONST int *s = reinterpret_cast<const MDX_CONST int *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = s[k];
    return v;
}
template <typename T>
__device__ inline T ldv(const MDX_GLOBAL T *p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized types only");
    T v;
    int *d = reinterpret_cast<int *>(&v);
    const MDX_GLOBAL int *s = reinterpret_cast<const MDX_GLOBAL int *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = s[k];
    return v;
}
template <typename T>
__device__ inline void stv(MDX_GLOBAL T *p, const T &v) {
    static_assert(sizeof(T) % 4 == 0, "word-sized types only");
    const int *s = reinterpret_cast<const int *>(&v);
    MDX_GLOBAL int *d = reinterpret_cast<MDX_GLOBAL int *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = s[k];
}

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

// The same on a global-memory pointer (a global atomic, not a flat one,
// so it does not also count against the scalar-memory wait counter).
__device__ inline void fx_add_g(MDX_GLOBAL unsigned long long *p, double v, double scale) {
    const long long q = __double2ll_rn(v * scale);
    __hip_atomic_fetch_add(p, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------
// Argument structs of the force-call kernels (one per system).

struct PackArgs {
    int nclu;
    const int *clu_atom;
    const double *anchor;
    const double *x, *y, *z, *q;
    Box box;
    float4 *xq;
    // molecular centres (skipped when user_com)
    int nmol, user_com;
    const int *mstart;
    const int *comgrp;           // molecule groups [comgrp[g], comgrp[g+1]) of the COM kernel
    const double *mass;
    double *cx, *cy, *cz;
};

struct PairArgs {
    int nclu;
    const PairEntry *list;       // per i-cluster: nshort entries with a short-class pair from
    const int *nlist;            //   the front, the other nlist - nshort from the back
    const int *nshort;
    int list_cap;
    const float4 *xq;            // per slot: x, y, z relative to anchor; q (FP32)
    const int2 *aux;             // per slot: LJ type | ee << 16; q - (float)q as float bits
    const double *q;             // per atom charge (FP64)
    float coulf_f;               // coulomb factor (FP32, forces)
    double coulf;                // coulomb factor (FP64, energies)
    const float2 *lj;            // nlj x nlj: (6 A6, 12 B12)
    const double2 *lj64;         // nlj x nlj: (A6, B12) in FP64 (k_pairs64)
    int nlj;
    const int *clu_atom;
    float alpha;
    float rc2;                   // evaluation cutoff^2 (Verlet) or +inf (reference)
    // Verlet lists: pairs whose FP32 r^2 lies within `band` of rc2 take the
    // cutoff test again in FP64 with the reference arithmetic (CHCNB's
    // one-shift minimum image, forces.f:1841-1868, and <= against R_cut^2),
    // so the set of evaluated pairs is that of a list built at the current
    // positions; band = 0 for reference lists (no cutoff test)
    float band;
    float rlo, rhi;              // rc2 - band, rc2 + band, rounded outwards
    double rc2d;
    const double *x, *y, *z;     // caller coordinates (FP64)
    Box box;
    float raex;
    double alphad, raexd;        // alpha and RAEX in FP64 (k_pairs64)
    int soft;                    // lambda-coupled types present
    float lx, ly, lz, ilx, ily, ilz;
    double *fs_x, *fs_y, *fs_z;  // short-class pair forces (per atom)
    double *fl_x, *fl_y, *fl_z;  // long-class pair forces (per atom)
    // j-atom reactions of the half list: k_pairs writes one record per list
    // entry at unit jdest[list slot] of rec (a unit is 24 floats: a
    // long-class entry's record is one unit, a front-segment entry's two);
    // the records naming cluster c as j-cluster are consecutive from unit
    // jstart[c]: jcnt[c].x front-segment records, then jcnt[c].y long-class
    // ones; k_jgather sums them into 64-bit fixed point (scale fscale), per
    // cluster slot 8 words: short x, y, z (fast forces H), long x, y, z (slow
    // forces G), 2 unused
    float *rec;
    const unsigned *jstart, *jdest;
    const int2 *jcnt;
    unsigned long long *jf;
    double fscale;
    double *epart;               // [nclu_cap][Q_N]
    // per-type breakdown (detail calls only)
    const int *atom_mol, *atom_moltype, *type_pair;
    int nmt, ntp, ndet;
    double *dpart;               // [nclu_cap][ndet]
};

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
    double *epart;
    int nmt, ntp, ndet;
    double *dpart;               // detail rows of this kernel's blocks
};

struct RecipArgs {
    int natoms;
    const double *x, *y, *z, *q;
    Box box;
    double4 *fq;                 // fractional coordinates and charge
    double4 *stepz;              // exp(2 pi i f_z), exp(16 pi i f_z)
    // phase tables, [index][tstride] (FP64 recurrences rounded to FP32,
    // rows zero-padded to tstride = a multiple of 64 atoms):
    // tx = q exp(2 pi i n_x f_x), n_x = 0..kx; ty = exp(2 pi i n_y f_y),
    // n_y = 0..ky; tz = exp(2 pi i j f_z), j = 0..mzp-1, stored split into
    // f16 hi/lo: per atom pair 16 B (C hi, D hi, C lo, D lo as half2)
    float2 *tx, *ty, *tz;
    int kx, ky, mzp;
    int tstride;                 // table row stride (atoms, a multiple of 64)
    int chunk, nchunk, nkv;
    double alpha, coulf;
    int nslice, tiles_per_slice, natom_blocks;
    double *rpart;               // [nslice][3][natoms]
    double *epart;               // rows of k_sk_reduce
    // (n_x, n_y) rows with +-n_z pairs: frows in (n_x, n_y) order (force
    // pass), srows the same rows sorted by pair count (structure factors,
    // groups of 64)
    int nfrows, npk, nsgroups;
    const FRow *frows;
    const FRow *srows;           // pad0: index of the row in frows
    const int4 *fpk;             // per pair entry: k(+j), k(-j), j, frow
    double4 *ppart;              // [nchunk][npk]: partial U_j, V_j (nchunk a multiple of 8)
    double2 *s0part;             // [nchunk][nfrows]: partial S(n_z = 0)
    float2 *fa0;                 // per row: A of n_z = 0
    // force pass: tiles of 4 rows and their matrix operand (frows: pad1 =
    // woff of the row's tile, pad2 = row within the tile)
    int ntiles;
    const FTile *ftiles;
    uint16_t *wm;                // f16 hi/lo operand, kForceStepH halves per step
};

struct ScatterArgs {
    const double *x, *y, *z;
    Box box;
    double fscale;
    unsigned long long *fx, *fy, *fz;
    double *epart;
};

struct EtermArgs {
    int np;
    const int2 *pairs;
    const double *q;
    double coulf, alpha;
    ScatterArgs s;
};

// tv (every bonded kernel): per term (value, energy) -- the bond length,
// bend angle, |dihedral| or, for STRBND, each record's three bonds -- for
// the per-template sums of k_term_reduce; written on every call.
struct BondArgs {
    int n;
    const int *bi, *bj, *bpar;
    const int *morse;
    const double *r0, *k, *d, *rho;
    ScatterArgs s;
    double2 *tv;                   // [n]
};

// Flexible SPC water (STRBND, forces.f:530-800): one molecule per record,
// atoms O, H1, H2 and 12 parameters (R_OH1, R_OH2, R_HH, A, B, C, rho1,
// rho2, E, G, IPOT, LMOVE).
struct SpcArgs {
    int n;
    const int *atom;
    const double *par;
    ScatterArgs s;
    double2 *tv;                   // [3 n]: O-H1, O-H2, H-H
};

struct AngleArgs {
    int n;
    const int *ai, *aj, *ak, *apar;
    const double *theta0, *k;
    ScatterArgs s;
    double2 *tv;                   // [n]
};

struct TorsArgs {
    int n;
    const int *ti, *tj, *tk, *tl, *tpar;
    const int *kind, *mult;
    const double *k, *delta, *k1, *k2, *k3, *k4, *k5;
    ScatterArgs s;
    double2 *tv;                   // [n]
};

// Per-template sums of the bonded terms' (value, energy) pairs (BR/BE,
// AR/AE, TR/TE): template m sums tv[slot[p]] * w[slot[p]] over p in
// [start[m], start[m + 1]) (slots in ascending order) into out[2 m],
// out[2 m + 1].
struct TermArgs {
    int ntm;
    const int *start, *slot;
    const double *w;               // 1 / NSPEC of the term's molecule type
    const double2 *tv;
    double *out;                   // [2 ntm]
};

// Path-integral ring: harmonic springs to the neighbouring beads
// (PIFORCES, pimd.f:321-394), into the fast fixed-point forces.
struct SpringArgs {
    int n;
    const double *x, *y, *z;               // this bead
    const double *xp, *yp, *zp;            // bead + 1
    const double *xm, *ym, *zm;            // bead - 1
    const double *mass;
    double beadfac, nbeads;
    Box box;
    double fscale;
    unsigned long long *fx, *fy, *fz;
    double *epart;
};

struct CombineArgs {
    int n;
    double scale;                  // 1 / P for path-integral beads, else 1
    const double *fsx, *fsy, *fsz; // pair short
    const double *flx, *fly, *flz; // pair long
    const double *recip;           // [nslice][3][n]
    int nslice;
    // fixed-point accumulators, cleared again once read (so the next call
    // starts from zeros without a fill kernel)
    unsigned long long *gx, *gy, *gz;    // slow scatter (ETERMS, long specials)
    unsigned long long *hlx, *hly, *hlz; // LOCALF specials
    unsigned long long *hbx, *hby, *hbz; // bonded
    unsigned long long *spx, *spy, *spz; // bead springs (unscaled) or null
    unsigned long long *jf;              // pair j reactions [slot][8]
    int clear_jf;                        // jf summed by atomics (FP64 mode): clear it too
    const int *atom_slot;
    double inv_scale;
    const double *x, *y, *z;
    const int *atom_mol;
    const double *cx, *cy, *cz;
    double *ox, *oy, *oz;          // F of the group
    double *Gx, *Gy, *Gz;          // slow (optional)
    double *Hx, *Hy, *Hz;          // fast (optional)
    double *epart;
};

struct DetailArgs {
    int nrows, ndet;
    const double *dpart;           // [nrows][ndet]
    double *out;                   // [ndet]
};

struct ReduceArgs {
    int nrows, nblk;               // nblk = ceil(nrows / kRedRows) blocks
    double *epart;                 // [nrows][Q_N], cleared once read
    double *out;                   // Q_N
    double *bpart;                 // [nblk][Q_N] block sums
};

// Device-resident leap-frog step (SULTAN, mdstep.f:29-260) of one system.
enum MdState { MD_SC = 0, MD_EKIN, MD_TEMP, MD_SHAKE_FAIL, MD_STEP, MD_NSTATE };
constexpr int kTrace = 5;          // per step: Epot, Ekin, T, SC, pair interactions

struct MdArgs {
    int n, nmol;
    double *x, *y, *z, *vx, *vy, *vz;
    double *ox, *oy, *oz;          // constrained positions at t (OX, mdstep.f:82-89)
    double *ux, *uy, *uz;          // unconstrained positions at t + dt
    const double *fx, *fy, *fz;    // total force F at t
    const double *gx, *gy, *gz;    // slow forces G (multiple time step)
    const double *hx, *hy, *hz;    // fast forces H
    const double *minv, *mass;     // 1/m and m per atom
    // dynamical mass m * mfac (imfac = 1 / mfac): 1, or 1 / P for the beads
    // of a path-integral ring (MASSD = MASS / NBEADS, setup.f:149)
    double mfac, imfac;
    const int *mstart;             // atoms of molecule m: [mstart[m], mstart[m+1])
    const int *atom_mol;           // molecule of every atom
    const int *comgrp;             // molecule groups of k_md_getcom
    const int *cstart;             // constraints of molecule m
    const int2 *cons;              // constrained atom pairs
    const double *cons_len2;       // RB^2
    double dt, accel, tol2, bdtol;
    double dts;                    // thermostat step: dt (leap-frog), dt / 2 (multiple time step)
    double dti;                    // short step dt / nfreq (multiple time step)
    int maxit, nvt;
    Box box;
    double ekin_factor, boltz, ndof, t0, dqt;
    double *state;                 // MdState
    double *kepart;                // per-block m v^2 partials
    int nke_blocks;
    const double *energy;          // Q_N energy slots of the last force call
    double spei;
    double *trace;                 // [step][kTrace] (step = state[MD_STEP])
    int trace_cap;
};

// Path-integral bead thermostat (BEADNOSE, pimd.f:402-441): one
// Nose-Hoover chain per atom, shared by its beads, driven by the atom's
// kinetic temperature averaged over the beads.
struct PimdArgs {
    int n, nbeads, nchain;
    double *const *v;              // [nbeads] -> the bead's velocities (x, y, z blocks of n)
    const double *mass;
    double *xi;                    // [nchain][n] chain variables (1 / time)
    double dti;                    // short step
    double inv_tau2;               // DQQT = 1 / BTAU^2
    double fnh;                    // DQQT / (3 nbeads): target of chain links 2..
    double t0;                     // TRTEMP
    double conv;                   // m v^2 -> bead-averaged kinetic temperature (bead mass m / nbeads):
                                   // ekin_factor / (3 k_B nbeads^2)
    // centroid wrapping (GETCOM with PI_COM, service.f:523-533)
    int nmol;
    const int *mstart;             // atoms of molecule m: [mstart[m], mstart[m+1])
    double *const *x, *const *y, *const *z;   // [nbeads] -> the bead's coordinates
    Box box;
};

}  // namespace mdx
