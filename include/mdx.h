/*
 * mdx.h -- C ABI of the MI355X force engine for MDynaMix.
 *
 * The engine replaces the force path of the reference program:
 *   CHCNB(LRDF)     neighbour lists                     forces.f:1813-2205
 *   ALLFORCE        all forces (SHAKE / leap-frog step)  forces.f:77-100
 *   FFORCES(LAVER)  fast forces (r-RESPA inner step)     forces.f:36-51
 *   SLFORCES        slow forces (r-RESPA outer step)     forces.f:55-73
 * together with what they call: LOCALF (813), FORCES (1014), FURIR (1217),
 * INIFUR (1356), ETERMS (1482), GETBND (113), GETANG (216), GETTRS (323)
 * and the reductions of CM_ADDLA/CM_ADDLF/CM_ADDF (scalar.f:28-116).
 *
 * The reference routines take no arguments and share state through COMMON
 * blocks (common_inc.f:133-249).  Here the same state crosses the boundary
 * as plain arrays: the topology and parameters once (mdx_create), the
 * coordinates every step (mdx_set_positions), and forces plus every
 * energy/virial scalar the integrators and averages read come back from
 * mdx_compute.  All calls return an mdx_status; nothing aborts the process
 * (the reference calls FINAL/stop instead, forces.f:1414-1419,2188-2194).
 *
 * Units are the caller's: coulomb_factor and the Lennard-Jones tables fix
 * the energy unit, coordinates fix the length unit, and forces come back in
 * energy/length.  The Fortran shim passes the reference's internal units
 * (COULF, A6LJ...), the Python host side passes kJ/mol, Angstrom, e.
 *
 * Indices are 0-based except in the exported reference-layout lists, which
 * keep the reference's 1-based signed form (negative first index marks a
 * 1-4 pair, forces.f:1956).
 *
 * Threading: one host thread per handle; a handle owns one HIP stream.
 */
#ifndef MDX_H
#define MDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mdx_status {
    MDX_OK = 0,
    MDX_ERR_ARG = 1,            /* invalid argument or unsupported option      */
    MDX_ERR_LIST_OVERFLOW = 2,  /* pair-list capacity exceeded (forces.f:2188)  */
    MDX_ERR_NKV = 3,            /* too many k-vectors (forces.f:1414)           */
    MDX_ERR_HIP = 4,            /* HIP runtime error                            */
    MDX_ERR_NOMEM = 5,          /* device allocation failed                     */
    MDX_ERR_STATE = 6,          /* call order (e.g. compute before a list)      */
    MDX_ERR_LIMIT = 7           /* beyond the fixed scratch of a call that
                                   allocates nothing                            */
} mdx_status;

/* Neighbour-list semantics. */
enum {
    /* CHCNB semantics: lists at R_cut with no skin; every listed pair is
     * evaluated until the next rebuild with no cutoff test (forces.f:820,
     * 1024); pairs keep the short/long class they had when listed. */
    MDX_LIST_REFERENCE = 0,
    /* Verlet list at R_cut + skin; pairs beyond R_cut at evaluation time
     * are skipped.  The short/long class is still fixed at build time. */
    MDX_LIST_VERLET = 1
};

/* Arithmetic of the force call (mdx_params.precision). */
enum {
    /* FP32 pair arithmetic from anchor-relative coordinates with FP64 and
     * 64-bit fixed-point sums; the reciprocal sum's phase products on the
     * matrix cores (f16 hi/lo pairs) with FP64 sums; FP64 bonded terms,
     * exclusion corrections, energies and virials. */
    MDX_PRECISION_MIXED = 0,
    /* FP64 throughout: the pair terms from the FP64 coordinates with the
     * reference's one-shift minimum image (forces.f:1078-1086), FURIR in
     * FP64 (mdx_furir_device's kernels).  The same list, classes and
     * outputs as MDX_PRECISION_MIXED; the accuracy reference of the
     * engine on the GPU, slower. */
    MDX_PRECISION_FP64 = 1
};

/* Force groups of mdx_compute. */
enum {
    MDX_FORCES_ALL = 0,   /* ALLFORCE: fast + slow, F = G + H                    */
    MDX_FORCES_FAST = 1,  /* FFORCES: short-list pairs + bonded terms -> H       */
    MDX_FORCES_SLOW = 2   /* SLFORCES: reciprocal + exclusions + long pairs -> G */
};

/* Optional work in mdx_compute. */
enum {
    MDX_WANT_DETAIL = 1,  /* per type-pair / per type energy breakdown
                             (POTE1, POTL1, PES141, PSR141, SPEE)             */
    MDX_WANT_TERMS = 2    /* per-template bonded sums BR/BE, AR/AE, TR/TE
                             (mdx_get_terms; fast-group calls only)          */
};

#define MDX_MAX_MOLTYPES 16
#define MDX_MAX_TYPEPAIRS (MDX_MAX_MOLTYPES * (MDX_MAX_MOLTYPES + 1) / 2 + 1)

/*
 * Topology and force-field tables (COMMON /ADDREFS/, /SITTES/, /SITSITM/,
 * /LISTNB/ INBB, /BONBON/, /ANGANG/, /TORTOR/, /SOMMOL/ C14*).
 * Atoms are ordered molecule type by molecule type, molecule by molecule,
 * as in the reference (setup.f:633-694).  All arrays are caller-owned and
 * only read during mdx_create.
 */
typedef struct mdx_topology {
    int32_t natoms;
    int32_t nmols;
    int32_t nmoltypes;            /* NTYPES (<= MDX_MAX_MOLTYPES)            */
    int32_t nljtypes;             /* NNBT                                    */
    int32_t ntypepairs;           /* MOLINT                                  */

    /* per atom */
    const int32_t *atom_moltype;  /* ITYPE - 1                               */
    const int32_t *atom_mol;      /* NNUM - 1                                */
    const int32_t *atom_ljtype;   /* INBT(NSITE) - 1                         */
    const double  *atom_charge;   /* Q: charge of intermolecular terms       */
    const double  *atom_charge0;  /* CHARGE(NSITE): unscaled site charge     */
    const double  *atom_mass;     /* MASS(NSITE), for molecular COMs         */

    /* per molecule type */
    const int32_t *type_ee;       /* IEE: 1 normal, 0 lambda-coupled type    */
    const int32_t *type_move;     /* LMOVE                                   */
    const double  *type_c14el;    /* C14EL                                   */
    const double  *type_c14lj;    /* C14LJ                                   */
    const int32_t *type_pair;     /* MDX(ITYP,JTYP) - 1, nmoltypes^2         */

    /* Lennard-Jones tables, nljtypes^2 row-major: E = (A6 + B12 r^-6) r^-6 */
    const double *a6, *b12;       /* A6LJ,  B12LJ   (intermolecular)         */
    const double *a6_0, *b12_0;   /* A6LJ0, B12LJ0  (intramolecular)         */
    const double *a6_14, *b12_14; /* A6_14, B12_14  (1-4 pairs)              */

    /* bound-atom lists NNBB/INBB (setup.f:753-888) in CSR form:
     * neighbours of atom i are excl_atom[excl_start[i] .. excl_start[i+1]),
     * excl_is14 marks the 1-4 entries (negative INBB entries).            */
    const int32_t *excl_start;
    const int32_t *excl_atom;
    const int8_t  *excl_is14;

    /* bonds (GETBND): E = K (r-r0)^2, or Morse when morse != 0 and r-r0<=r0 */
    int32_t nbonds;
    const int32_t *bond_i, *bond_j, *bond_param;
    int32_t nbondparams;
    const int32_t *bp_morse;      /* ID == 1                                 */
    const double  *bp_r0, *bp_k, *bp_d, *bp_rho;

    /* angles (GETANG): E = K (theta-theta0)^2, theta0 in radians */
    int32_t nangles;
    const int32_t *angle_i, *angle_j, *angle_k, *angle_param;
    int32_t nangleparams;
    const double  *ap_theta0, *ap_k;

    /* dihedrals (GETTRS) */
    int32_t ntors;
    const int32_t *tors_i, *tors_j, *tors_k, *tors_l, *tors_param;
    int32_t ntorsparams;
    const int32_t *tp_kind;       /* ITORS: 0 AMBER, 1 MM3, 5 cos^n, -1 improper */
    const int32_t *tp_mult;       /* NMUL                                    */
    const double  *tp_k;          /* FT                                      */
    const double  *tp_delta;      /* RT, radians                             */
    const double  *tp_k1, *tp_k2, *tp_k3, *tp_k4, *tp_k5;   /* FT1..FT5      */
    /* flexible SPC water of the "fSPC" kind (STRBND, forces.f:530-800,
       called from GETBND for types with IPOT = 1 or 2, whose bonds and
       angles are then not listed, setup.f:297): one record per molecule,
       atoms O, H1, H2 and 12 parameters R_OH1, R_OH2, R_HH (RB of the
       type's bonds 1-3), A, B (DB of bonds 1, 2 for IPOT 2, FB for 1),
       C (FB of bond 3), rho1, rho2 (RO of bonds 1, 2), E, G (the fixed
       cross constants), IPOT, LMOVE; energies in the units of the other
       bonded parameters.  nspc = 0 when no such type exists. */
    int32_t nspc;
    const int32_t *spc_atom;      /* [3 nspc] */
    const double  *spc_par;       /* [12 nspc] */
    /* [nspc] or NULL: bond template (parameter index, IADB - 1) of each
       record's R_OH1 bond; R_OH2 and R_HH are the next two templates.
       STRBND adds its bond lengths and energies there (BR/BE,
       forces.f:655-663); NULL leaves them out of mdx_get_terms. */
    const int32_t *spc_bond;
} mdx_topology;

/* Run parameters (COMMON /SIZES/, /ENERGY/ COULF, /MDEE/ RAEX). */
typedef struct mdx_params {
    double box[3];                /* BOXL, BOYL, BOZL (rectangular cell)     */
    double rcut;                  /* RCUTT                                   */
    double rshort;                /* SHORT                                   */
    double skin;                  /* Verlet skin (MDX_LIST_VERLET only)      */
    double alpha;                 /* ALPHAD (Ewald); 0 disables Ewald        */
    double fexp;                  /* FEXP: k-space cut |k|^2 <= 4 alpha^2 FEXP */
    double coulomb_factor;        /* COULF                                   */
    int32_t list_mode;            /* MDX_LIST_REFERENCE / MDX_LIST_VERLET    */
    int32_t ee_enabled;           /* LMDEE                                   */
    double ee_raex;               /* RAEX: soft-core shift of EE pairs       */
    double ee_self_scale;         /* EC(ME): charge scale of EE types in the
                                     Ewald self term (forces.f:1434-1435)   */
    int32_t precision;            /* MDX_PRECISION_MIXED / MDX_PRECISION_FP64 */
} mdx_params;

/* Energies and virials of one force call (COMMON /ENERGY/, /VIRRR/,
 * /SOMMOL/ as left by ZEROFS + the force routines + CM_ADD*). */
typedef struct mdx_energies {
    double pe1, pe2;              /* LJ energy, long / short list            */
    double pels1, pels2;          /* electrostatic energy (slow incl. k-space) / short */
    double spe;                   /* Ewald self + exclusion correction       */
    double pint;                  /* bonded energy                           */
    double vir1, vir2, virb;      /* LJ virials (long/short), bond virial    */
    double virx, viry, virz;      /* slow virial projections                 */
    double virfx, virfy, virfz;   /* fast virial projections                 */
    double wirs, wirss;           /* molecular virial corrections            */
    int64_t npairs_short;         /* pairs evaluated in this call            */
    int64_t npairs_long;
    /* MDX_WANT_DETAIL only */
    double pote1[MDX_MAX_TYPEPAIRS];
    double potl1[MDX_MAX_TYPEPAIRS];
    double pes141[MDX_MAX_MOLTYPES];
    double psr141[MDX_MAX_MOLTYPES];
    double spee[MDX_MAX_MOLTYPES];
    /* path-integral batches (mdx_batch_set_ring): spring energy QEK and
     * spring virial VIRPI of this bead (PIFORCES, pimd.f:321-394) */
    double qek, virpi;
} mdx_energies;

typedef struct mdx_engine mdx_engine;

/* Library and device. */
const char *mdx_version(void);
mdx_status mdx_device_info(int32_t *n_cu, int32_t *clock_mhz, char *name, int32_t name_len);

/* Handle life cycle.  mdx_create copies topology and parameters to the
 * device, enumerates the k-vectors (INIFUR) and computes the self terms. */
mdx_status mdx_create(const mdx_topology *topo, const mdx_params *params, mdx_engine **out);
mdx_status mdx_destroy(mdx_engine *h);
const char *mdx_last_error(const mdx_engine *h);

/* Coordinates: FP64 SoA, caller-owned host arrays of natoms (SINHR/GETCOM
 * output).  The device-pointer variant takes arrays already in HBM.       */
mdx_status mdx_set_positions(mdx_engine *h, const double *x, const double *y, const double *z);
mdx_status mdx_set_positions_device(mdx_engine *h, const double *dx, const double *dy, const double *dz);
/* Box change (NPT); re-enumerates k-vectors like FURIR when an edge changed
 * by more than 5 % since the last enumeration (forces.f:1250-1265). */
mdx_status mdx_set_box(mdx_engine *h, const double box[3]);
/* Molecular centres of mass used by the molecular virial (X, Y, Z of
 * COMMON /COMCOM/).  NULL restores the default: COMs from atom masses. */
mdx_status mdx_set_molecule_com(mdx_engine *h, const double *x, const double *y, const double *z);

/* CHCNB: rebuild the neighbour lists from the current coordinates. */
mdx_status mdx_build_list(mdx_engine *h);
/* Pair counts of the last build (MBSH, MBLN). */
mdx_status mdx_list_counts(mdx_engine *h, int64_t *n_short, int64_t *n_long);
/* Cluster-pair entries of the last build (8 x 8 atom pairs each): all of
 * them, and those holding a short-class pair.  No reference counterpart
 * (the reference lists atom pairs); it gives the pair kernel's lane count. */
mdx_status mdx_list_entries(mdx_engine *h, int64_t *entries, int64_t *front_entries);
/* Export the lists in the reference layout and order (NBS1/NBS2, NBL1/NBL2:
 * 1-based, the larger index of a normal pair first, 1-4 pairs as
 * (-outer, inner) in the short list).  Capacities in pairs. */
mdx_status mdx_get_list(mdx_engine *h, int32_t *nbs1, int32_t *nbs2, int64_t cap_short,
                        int32_t *nbl1, int32_t *nbl2, int64_t cap_long);

/* One force call.  group is MDX_FORCES_*.  Output arrays are caller-owned
 * host FP64 arrays of natoms; any may be NULL.  f receives the forces of
 * the group (ALL: G + H), g and h the slow and fast parts when non-NULL. */
mdx_status mdx_compute(mdx_engine *h, int32_t group, int32_t want,
                       double *fx, double *fy, double *fz,
                       double *gx, double *gy, double *gz,
                       double *hx, double *hy, double *hz,
                       mdx_energies *e);
/* Same call with device output arrays of natoms (forces stay in HBM). */
mdx_status mdx_compute_device(mdx_engine *h, int32_t group, int32_t want,
                              double *dfx, double *dfy, double *dfz, mdx_energies *e);

/* Per-template bonded sums of the last force call made with MDX_WANT_TERMS
 * that evaluated the fast group (MDX_FORCES_ALL or MDX_FORCES_FAST), as
 * GETBND, STRBND, GETANG and GETTRS leave them for CM_ADDF / CM_ADDLA
 * (forces.f:119-182, 655-663, 232-273, 336-447; scalar.f:42-51, 95-106):
 * over the terms of template m, each divided by NSPEC of its molecule type,
 *   br[m] bond lengths, be[m] bond energies (STRBND records included),
 *   ar[m] bend angles (radians), ae[m] their energies,
 *   tr[m] |dihedral angle| (radians), te[m] their energies
 * (dihedrals with collinear bonds contribute nothing).  br/be hold
 * nbondparams values, ar/ae nangleparams, tr/te ntorsparams; any may be
 * NULL.  Zeros before the first such call.  Sums in a fixed order. */
mdx_status mdx_get_terms(mdx_engine *h, double *br, double *be, double *ar, double *ae, double *tr, double *te);

/* Expanded-ensemble energies at trial coupling parameters: the sums DIFEN
 * and DIFFUR form for a move between subensembles (mdee.f:265-534), as
 * functions of lambda, at the current coordinates and list, in FP64, for
 * nlam <= MDX_EE_MAX_LAMBDA values of lambda at once.  The reference's
 * energies toward subensemble ME1 are EEN = e_pair(EC(ME1)),
 * EEO = e_pair(EC(ME)), EFN = e_recip(EC(ME1)), EFO = e_recip(EC(ME)).
 * e_pair: over the listed pairs with an atom of a lambda-coupled type
 * (type_ee = 0) in different molecules (reference lists: no cutoff test;
 * Verlet lists: r <= R_cut), s = lambda^PW1, shift = BEE (1 - lambda^PW2):
 *   s^(coupled atoms) q0_i q0_j COULF erfc(alpha r) / (r + shift)
 *   + s (A6_0 + B12_0 (r + shift)^-6) (r + shift)^-6.
 * e_recip: sum_k AKC |S(k)|^2 with the coupled atoms' charges s q0, minus
 * their self terms COULF (s q0)^2 alpha / sqrt(pi), plus over the pairs
 * inside coupled molecules s^2 COULF q0_i q0_j (erfc(alpha r) - 1) / r.
 * Either output may be NULL.  Synchronous; needs a built list. */
#define MDX_EE_MAX_LAMBDA 32
typedef struct mdx_ee_params {
    double pw1, pw2, bee;         /* EE_power PW1 PW2, EE_offset BEE      */
} mdx_ee_params;
mdx_status mdx_ee_energies(mdx_engine *h, const mdx_ee_params *p, int32_t nlam, const double *lambda,
                           double *e_pair, double *e_recip);

/* Forces of the last call (F, G, H) to caller-owned host arrays (any may be
 * NULL), e.g. after a batched call. */
mdx_status mdx_get_forces(mdx_engine *h, double *fx, double *fy, double *fz, double *gx, double *gy, double *gz,
                          double *hx, double *hy, double *hz);
/* Per-stage HIP-event timing of mdx_compute (on by default; off replays the
 * call from a captured hipGraph). */
mdx_status mdx_set_timing(mdx_engine *h, int32_t on);
/* Autotuner: times the reciprocal-space kernels of this system over a set
 * of launch shapes (atoms per structure-factor block; row slices of the
 * force pass), then the pair stage over cluster geometries (the width of
 * the x-y columns whose z-sorted atoms form the 8-atom clusters: every
 * column count along x between 0.75 and 1.3 times the cube root of 8
 * atoms' volume, scored by the pair stage plus a tenth of a list build)
 * and over the pair kernel's occupancy (five or six wavefronts per SIMD),
 * on the current coordinates, and
 * keeps the fastest; the list is rebuilt at the current coordinates with
 * the chosen grid.  The choice is remembered for the process by size
 * class (atoms, k vectors) and used by later mdx_create calls of that
 * class.  Shape and grid change only how the sums are split into partial
 * sums (results agree to rounding; each choice is deterministic).  Needs a
 * built list.  No counterpart in the reference. */
mdx_status mdx_autotune(mdx_engine *h, int32_t *chunk, int32_t *nslice);
/* A fixed cluster grid instead of the tuned one: column width = scale
 * (0.5 to 2) times the cube root of 8 atoms' volume; the list is rebuilt
 * at the current coordinates when one was built.  No counterpart in the
 * reference. */
mdx_status mdx_set_cluster_grid(mdx_engine *h, double scale);
/* The cluster grid in use: column width (A, along x) and the column counts
 * along x and y.  No counterpart in the reference. */
mdx_status mdx_cluster_grid(const mdx_engine *h, double *column_width, int32_t *ncx, int32_t *ncy);
/* Wavefronts per SIMD of the plain pair kernel (5 or 6; the kernel's
 * arithmetic is the same, only its register budget differs): set (a fixed
 * choice for profiling runs) and read.  No counterpart in the reference. */
mdx_status mdx_set_pair_waves(mdx_engine *h, int32_t waves);
mdx_status mdx_pair_waves(const mdx_engine *h, int32_t *waves);
/* The k vectors of the reciprocal sum in INIFUR order (forces.f:1375-1427,
 * rectangular cell): integer triples (n_x, n_y, n_z) with k = 2 pi n / L,
 * i.e. RKX = 2 pi n_x as INIFUR stores it.  *nkv receives the count; the
 * arrays (cap entries, may be NULL when cap is 0) the vectors.  A single
 * molecule has none (FURIR returns at once, forces.f:1266). */
mdx_status mdx_get_kvectors(const mdx_engine *h, int32_t *nx, int32_t *ny, int32_t *nz, int64_t cap,
                            int64_t *nkv);
/* The same enumeration from parameters alone (box, alpha, fexp), on the
 * host, without a device or a handle. */
mdx_status mdx_enumerate_kvectors(const mdx_params *p, int32_t *nx, int32_t *ny, int32_t *nz, int64_t cap,
                                  int64_t *nkv);
/*
 * FURIR in FP64 on device arrays, without a handle: the reciprocal Ewald
 * forces (forces.f:1320-1326), the energy QPE (forces.f:1317-1318) and the
 * virial projections (forces.f:1328-1335) of natoms charges over the k set
 * INIFUR enumerates from p (box, alpha, fexp; rectangular cell,
 * forces.f:1375-1427).  x, y, z, q and the outputs fx, fy, fz (forces,
 * written) and energy4 (QPE, VIRX, VIRY, VIRZ) are device arrays; the
 * caller's workspace (device memory of mdx_furir_workspace_bytes bytes)
 * holds the k set, the structure factors and the partial sums.  Work is
 * queued on `stream` (a hipStream_t; NULL: the default stream) and not
 * waited for.  Every sum runs in a fixed order, so repeated calls give the
 * same bits.  FURIR's single-molecule exit (forces.f:1266) is the
 * caller's: this routine always sums.  Replaces FURIR's arithmetic
 * (forces.f:1271-1335) for callers holding coordinates in HBM; the engine
 * uses it in its FP64 precision mode.
 */
mdx_status mdx_furir_workspace_bytes(const mdx_params *p, int32_t natoms, int64_t *bytes);
mdx_status mdx_furir_device(const mdx_params *p, int32_t natoms, const double *x, const double *y, const double *z,
                            const double *q, double *fx, double *fy, double *fz, double *energy4, void *workspace,
                            int64_t workspace_bytes, void *stream);
/*
 * mdx_furir_device without the workspace argument, for callers that cannot
 * allocate device memory (code under hipGraph stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The k set, structure factors and partial sums live in a
 * static device array of MDX_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; calls whose mdx_furir_workspace_bytes exceeds it return
 * MDX_ERR_LIMIT before any device work.  The call allocates nothing (no
 * hipMalloc, hipFree or allocating library routine).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two host
 * threads, or two streams in flight) race and must be serialised by the
 * caller; consecutive calls on one stream are safe.
 */
#define MDX_NOALLOC_SCRATCH_BYTES (64LL << 20)
mdx_status mdx_furir_device_noalloc(const mdx_params *p, int32_t natoms, const double *x, const double *y,
                                    const double *z, const double *q, double *fx, double *fy, double *fz,
                                    double *energy4, void *stream);
/* Set that shape directly (reproducible runs, profiling). */
mdx_status mdx_set_kspace_shape(mdx_engine *h, int32_t chunk, int32_t nslice);
/* Device memory held by a handle, in bytes (admission by HBM budget). */
mdx_status mdx_memory_bytes(const mdx_engine *h, int64_t *bytes);
/* Device memory mdx_create would allocate for this system (plus the MD
 * segment state when with_md), computed without allocating. */
mdx_status mdx_estimate_bytes(const mdx_topology *t, const mdx_params *p, int32_t with_md, int64_t *bytes);

/*
 * Batches: K systems (path-integral beads, expanded-ensemble lambda states,
 * independent jobs) evaluated by one set of fused launches per force call:
 * every kernel covers all K systems through a block map, and the sequence
 * is captured once in a hipGraph and replayed.  Systems keep their own
 * lists (mdx_build_list per system); a system may be in one batch at a time
 * besides its own mdx_compute.  No counterpart in the reference, which runs
 * one bead per MPI rank (pimd.f:18, input.f:651-659).
 */
typedef struct mdx_batch mdx_batch;
mdx_status mdx_batch_create(mdx_engine *const *systems, int32_t k, mdx_batch **out);
/* Path-integral ring over the K systems (beads of one system): harmonic
 * springs BEADFAC * m * |r_k - r_k+-1|^2 to both neighbours (PIFORCES,
 * pimd.f:321-394) added to the fast forces, physical forces scaled by 1/K
 * (pimd.f:61-69). */
mdx_status mdx_batch_set_ring(mdx_batch *b, double beadfac);
mdx_status mdx_batch_use_graph(mdx_batch *b, int32_t on);
/* One force call of group MDX_FORCES_* for every system; e: K entries or
 * NULL.  Forces stay in HBM; fetch them with mdx_get_forces. */
mdx_status mdx_batch_compute(mdx_batch *b, int32_t group, int32_t want, mdx_energies *e);
const char *mdx_batch_last_error(const mdx_batch *b);
mdx_status mdx_batch_destroy(mdx_batch *b);

/*
 * Device-resident MD segments: the leap-frog step of SULTAN
 * (mdstep.f:29-260) with SHEJK bond constraints (mdstep.f:1105-1283), the
 * Nose-Hoover thermostat of SCALING(3) (mdstep.f:559-589, 908-925) and
 * GETCOM molecule wrapping (service.f:463-563), all in FP64 on the device;
 * the list is rebuilt every rebuild_every steps (ICHNB) and the forces are
 * the ALLFORCE of this engine.  Units are the caller's: acceleration =
 * accel_factor * F / m, kinetic energy = ekin_factor * m v^2 / 2,
 * T = 2 E_kin / (ndof * boltz).
 */
typedef struct mdx_md_params {
    double dt;                /* TSTEP (time unit of the velocities)        */
    double accel_factor;
    double ekin_factor;
    double boltz;             /* k_B in energy units per K                  */
    double temperature;       /* TRTEMP                                     */
    double tau_t;             /* Nose-Hoover QT (time units); 0: NVE        */
    double sc0;               /* initial thermostat variable SC (1/time)    */
    double ndof;              /* degrees of freedom FNST                    */
    double shake_tol;         /* TOLER                                      */
    double shake_bdtol;       /* BDTOL                                      */
    int32_t shake_maxit;      /* MAXIT (2000)                               */
    int32_t rebuild_every;    /* ICHNB                                      */
    int32_t step0;            /* MSTEP before the first step (NSTTOT)       */
    int32_t ncons;            /* constrained bonds, grouped by molecule,    */
    const int32_t *cons_i;    /*   in the reference bond order (0-based)    */
    const int32_t *cons_j;
    const double *cons_len;   /* RB                                         */
    int32_t integrator;       /* MDX_MD_LEAPFROG, MDX_MD_RESPA, MDX_MD_PIMD */
    int32_t nfreq;            /* NFREQ: short steps per long step           */
} mdx_md_params;

/* Integrators of mdx_md_run (mdx_md_params.integrator):
 * MDX_MD_LEAPFROG  SULTAN (mdstep.f:29-260): leap-frog, SHEJK constraints,
 *                  SCALING(3), one force call (ALLFORCE) per step;
 * MDX_MD_RESPA     DOUBLE (mdstep.f:281-530): multiple time step, dt the
 *                  long step; per long step SCALING(1), a half kick with
 *                  the slow forces G, nfreq short steps (half kick with the
 *                  fast forces H, drift, GETCOM, fast forces -- slow and
 *                  fast at the last short step --, half kick), a half kick
 *                  with G, SCALING(2); the list is rebuilt before the force
 *                  call of short step nfreq/2 + 1 of every ICHNB-th long
 *                  step; no constraints (ncons must be 0). */
#define MDX_MD_LEAPFROG 0
#define MDX_MD_RESPA 1
/* MDX_MD_PIMD      PIMD (pimd.f:18-314): the beads of a path-integral ring
 *                  batch (mdx_batch_set_ring) advanced by the multiple-time-
 *                  step scheme of MDX_MD_RESPA with the forces of the ring
 *                  (physical forces / P, PIFORCES springs added to the fast
 *                  forces) and, in every short step between the fast half
 *                  kick and the drift, BEADNOSE's per-atom thermostat chains
 *                  (mdx_batch_pimd_init); every bead is initialised with
 *                  this integrator, the same nfreq and tau_t = 0 (the
 *                  reference's PI_thermo switches the Nose thermostat off,
 *                  input.f:660-664); run with mdx_batch_md_run. */
#define MDX_MD_PIMD 2

/* Start a segment from the current positions and the given velocities:
 * builds the list and evaluates the forces (SULTAN's CHCNB + ALLFORCE). */
mdx_status mdx_md_init(mdx_engine *h, const mdx_md_params *p, const double *vx, const double *vy,
                       const double *vz);
/* nsteps steps (long steps under MDX_MD_RESPA); trace (NULL or nsteps x 5
 * doubles): per step the potential
 * energy, kinetic energy, temperature, SC and the number of pair
 * interactions of the step's force call. */
mdx_status mdx_md_run(mdx_engine *h, int32_t nsteps, double *trace);
mdx_status mdx_md_get_state(mdx_engine *h, double *x, double *y, double *z, double *vx, double *vy, double *vz);
/* Constraint failures since mdx_md_init: molecules whose SHEJK sweeps hit
 * MAXIT without convergence, plus molecules in which a bond update's
 * projection on the old bond fell below BDTOL (one count per molecule and
 * step; the reference prints "NO CONVERGENCE" and
 * continues, mdstep.f:1264-1268; the device run continues too and counts
 * them here).  0 in a healthy run. */
mdx_status mdx_md_shake_failures(mdx_engine *h, int64_t *count);
/* nsteps steps of every system of a batch (each initialised by
 * mdx_md_init), one fused launch per kernel and step; trace: K x nsteps x 5
 * or NULL. */
mdx_status mdx_batch_md_run(mdx_batch *b, int32_t nsteps, double *trace);

/* Bead thermostat of a path-integral run ("PI_thermo T BTAU JNH",
 * input.f:660-666): one Nose-Hoover chain of nchain links per atom, shared
 * by the atom's beads, driven by its kinetic temperature averaged over the
 * beads (BEADNOSE, pimd.f:402-441); tau_bead 0: no bead thermostat (NVE).
 * Units as mdx_md_params (tau in its time unit). */
typedef struct mdx_pimd_params {
    double temperature;       /* TRTEMP                                     */
    double tau_bead;          /* BTAU                                       */
    int32_t nchain;           /* JNH (>= 1)                                 */
} mdx_pimd_params;
/* Make a ring batch a path-integral MD run (PIMD, pimd.f:18-314): every
 * bead initialised by mdx_md_init with MDX_MD_PIMD; the chains start at
 * zero and the ring forces are evaluated at the current positions
 * (pimd.f:46-69).  Then mdx_batch_md_run advances the beads. */
mdx_status mdx_batch_pimd_init(mdx_batch *b, const mdx_pimd_params *p);
/* The chain variables (1 / time), nchain x n (link-major). */
mdx_status mdx_batch_pimd_chains(mdx_batch *b, double *xi);

/*
 * Serving: MD-segment jobs admitted by an HBM budget and advanced together
 * (continuous batching).  A worker thread owns the device; jobs submitted
 * while others run join the batch as slots and memory free up.  No
 * counterpart in the reference (one job per program run).
 */
typedef struct mdx_job_desc {
    const mdx_topology *topology;     /* read at admission (mdx_create)     */
    const mdx_params *params;
    const mdx_md_params *md;
    const double *x, *y, *z;          /* start coordinates                  */
    const double *vx, *vy, *vz;       /* start velocities                   */
    int32_t nsteps;                   /* segment length                     */
    double *x_out, *y_out, *z_out;    /* final state (any may be NULL)      */
    double *vx_out, *vy_out, *vz_out;
} mdx_job_desc;                       /* every array stays valid until the
                                         job's result is polled            */
typedef struct mdx_job_result {
    int64_t id;
    int32_t status;                   /* mdx_status of the job              */
    int32_t nsteps;                   /* steps completed                    */
    double t_submit, t_admit, t_done; /* seconds on the server clock        */
    double epot, ekin, temperature;   /* after the last step                */
    int64_t pair_interactions;        /* over all of the job's force calls  */
    int64_t bytes;                    /* device memory estimate admitted    */
    char message[128];
} mdx_job_result;
typedef struct mdx_server mdx_server;
/* quantum: steps per scheduling round (admission/retirement granularity). */
mdx_status mdx_serve_create(int64_t hbm_budget_bytes, int32_t max_slots, int32_t quantum, mdx_server **out);
mdx_status mdx_serve_submit(mdx_server *s, const mdx_job_desc *job, int64_t *id);
/* Completed jobs (up to max), and the current active / queued counts. */
mdx_status mdx_serve_poll(mdx_server *s, mdx_job_result *out, int32_t max, int32_t *n, int32_t *active,
                          int32_t *queued);
/* Block until no job is queued or active (MDX_ERR_STATE on timeout). */
mdx_status mdx_serve_wait(mdx_server *s, double timeout_s);
mdx_status mdx_serve_stats(mdx_server *s, int64_t *rounds, double *busy_s, double *clock_s);
const char *mdx_serve_last_error(const mdx_server *s);
mdx_status mdx_serve_destroy(mdx_server *s);

/* Timing of the last mdx_build_list and mdx_compute per stage, in ms (HIP
 * events on the handle's stream): list build, prep (pack + COM), pairs,
 * special pairs, k-space prep, structure factors, S(k) reduction,
 * k-space forces, exclusions + bonded, force assembly + energy sums. */
mdx_status mdx_stage_times(mdx_engine *h, double *ms, int32_t n);

#ifdef __cplusplus
}
#endif
#endif /* MDX_H */
