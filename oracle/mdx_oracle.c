/*
 * mdx_oracle.c -- FP64 CPU restatement of the MDynaMix force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
 * MI355X engine: it is linked into oracle/liborc.so and called only by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The engine never calls it.
 *
 * It restates, in plain C double precision with FMA contraction off (built
 * with -ffp-contract=off, so every product and sum rounds as on an x86-64
 * gfortran build), the following reference routines:
 *
 *   orc_chcnb     CHCNB    forces.f:1813-2205  (rectangular cell, Ewald);
 *                 orc_chcnb_task is one rank's stripe of it
 *   orc_localf    LOCALF   forces.f:813-998
 *   orc_forces    FORCES   forces.f:1014-1202
 *   orc_inifur    INIFUR   forces.f:1356-1449
 *   orc_furir     FURIR    forces.f:1217-1339 (orc_furir_task: one rank's k stripe)
 *   orc_eterms    ETERMS   forces.f:1482-1676
 *   orc_strbnd    STRBND   forces.f:530-800 (fSPC water)
 *   orc_getbnd    GETBND   forces.f:113-204
 *   orc_getang    GETANG   forces.f:216-309
 *   orc_gettrs    GETTRS   forces.f:323-513
 *   orc_compute   ALLFORCE/FFORCES/SLFORCES forces.f:36-100 with ZEROFS
 *                 (service.f:397-453) and CM_ADDLA/CM_ADDLF
 *                 (scalar.f:59-116)
 *
 * Inputs are the same mdx_topology / mdx_params structs the engine takes
 * (include/mdx.h), 0-based; lists are in the reference's 1-based signed
 * layout.  How each function is pinned is stated in its comment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mdx.h"

/* prcm.h:27-32 */
#define A1 0.254829592
#define A2 (-0.284496736)
#define A3 1.421413741
#define A4 (-1.453152027)
#define A5 1.061405429
#define B1 0.3275911
#define PI_R 3.1415926536
#define TOTPI 1.128379169

/* Fortran SIGN(a,b): |a| with the sign of b (b == 0 -> +|a|). */
static inline double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* Inline minimum image of the pair loops: x and z use the x half-box in
 * LOCALF (forces.f:872-873) and the z half-box everywhere else
 * (forces.f:1077-1078, 1518-1519, 1841-1842); y as forces.f:898.        */
static inline void min_image(double *dx, double *dy, double *dz, const mdx_params *p, int localf_quirk) {
    const double hx = 0.5 * p->box[0], hy = 0.5 * p->box[1], hz = 0.5 * p->box[2];
    if (fabs(*dx) > hx) *dx = *dx - fsign(p->box[0], *dx);
    if (fabs(*dz) > (localf_quirk ? hx : hz)) *dz = *dz - fsign(p->box[2], *dz);
    if (fabs(*dy) > hy) *dy = *dy - fsign(p->box[1], *dy);
}

/* PBC (service.f:953-990), rectangular: used by the bonded routines. */
static inline void pbc(double *dx, double *dy, double *dz, const mdx_params *p) {
    min_image(dx, dy, dz, p, 0);
}

/* ------------------------------------------------------------------ */
/* CHCNB (forces.f:1813-2205), NUMTASK = 1, rectangular cell, Ewald
 * (LMNBL = .false., so RRM = RR2).  Outer atom I ascending; "even strip"
 * J = I-2, I-4, ... then "odd strip" J = I+1, I+3, ...  Same-molecule pairs
 * with RR2 <= 64 are checked against INBB(I,*): positive entries drop the
 * pair, negative ones put it in the short list as (-I, J).  Otherwise the
 * pair goes to the short list if RR2 <= SHORT2, else to the long list if
 * RR2 <= RSSQ; even-strip pairs are stored (I, J), odd-strip ones (J, I).
 * Pinned by: the reference's own lists dumped by oracle/ref (bit-exact
 * array equality in tests/test_oracle_golden.py).                      */
static int64_t chcnb_visit(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                           const double *z, int i, int j, int odd, int32_t *s1, int32_t *s2, int64_t cap_s,
                           int64_t *ms, int32_t *l1, int32_t *l2, int64_t cap_l, int64_t *ml) {
    const double short2 = p->rshort * p->rshort;
    const double rssq = p->rcut * p->rcut;
    double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
    min_image(&dx, &dy, &dz, p, 0);
    double rr2 = dx * dx + dy * dy;
    rr2 = rr2 + dz * dz;
    if (t->atom_mol[i] == t->atom_mol[j] && rr2 <= 64.0) {
        for (int k = t->excl_start[i]; k < t->excl_start[i + 1]; ++k) {
            if (t->excl_atom[k] != j) continue;
            if (!t->excl_is14[k]) return 0;             /* 1-2 / 1-3 */
            if (*ms < cap_s) { s1[*ms] = -(i + 1); s2[*ms] = j + 1; }
            (*ms)++;
            return 0;
        }
    }
    int a = odd ? j : i, b = odd ? i : j;
    if (rr2 <= short2) {
        if (*ms < cap_s) { s1[*ms] = a + 1; s2[*ms] = b + 1; }
        (*ms)++;
    } else if (rr2 <= rssq) {
        if (*ml < cap_l) { l1[*ml] = a + 1; l2[*ml] = b + 1; }
        (*ml)++;
    }
    return 0;
}

/* Task-striped variant: outer atoms I = NUMTASK-TASKID, +NUMTASK, ...
 * (1-based; forces.f:1825-1830), i.e. the share of one replicated-data
 * rank.  numtask = 1 is the serial CHCNB. */
int orc_chcnb_task(const mdx_topology *t, const mdx_params *p, const double *x, const double *y, const double *z,
                   int taskid, int numtask, int32_t *nbs1, int32_t *nbs2, int64_t cap_s, int64_t *mbsh,
                   int32_t *nbl1, int32_t *nbl2, int64_t cap_l, int64_t *mbln) {
    int64_t ms = 0, ml = 0;
    const int n = t->natoms;
    for (int i = numtask - taskid - 1; i < n; i += numtask) {
        for (int j = i - 2; j >= 0; j -= 2)
            chcnb_visit(t, p, x, y, z, i, j, 0, nbs1, nbs2, cap_s, &ms, nbl1, nbl2, cap_l, &ml);
        for (int j = i + 1; j < n; j += 2)
            chcnb_visit(t, p, x, y, z, i, j, 1, nbs1, nbs2, cap_s, &ms, nbl1, nbl2, cap_l, &ml);
    }
    *mbsh = ms;
    *mbln = ml;
    return (ms > cap_s || ml > cap_l) ? MDX_ERR_LIST_OVERFLOW : MDX_OK;
}

int orc_chcnb(const mdx_topology *t, const mdx_params *p, const double *x, const double *y, const double *z,
              int32_t *nbs1, int32_t *nbs2, int64_t cap_s, int64_t *mbsh,
              int32_t *nbl1, int32_t *nbl2, int64_t cap_l, int64_t *mbln) {
    return orc_chcnb_task(t, p, x, y, z, 0, 1, nbs1, nbs2, cap_s, mbsh, nbl1, nbl2, cap_l, mbln);
}

/* ------------------------------------------------------------------ */
typedef struct orc_acc {
    double *fx, *fy, *fz;         /* target force array (G or H) */
    mdx_energies *e;
} orc_acc;

/* Shared parameter selection of LOCALF/FORCES (forces.f:835-862,
 * 1039-1066). */
static void pair_params(const mdx_topology *t, const mdx_params *p, int isp0neg, int isp, int jsp,
                        double *qij, double *a6, double *b12, double *rae) {
    const int ityp = t->atom_moltype[isp], jtyp = t->atom_moltype[jsp];
    const int isn = t->atom_ljtype[isp], jsn = t->atom_ljtype[jsp];
    const int nlj = t->nljtypes;
    *rae = 0.0;
    if (t->atom_mol[isp] == t->atom_mol[jsp]) {
        if (isp0neg) {
            *qij = t->atom_charge0[isp] * t->atom_charge0[jsp] * p->coulomb_factor * t->type_c14el[ityp];
            *a6 = t->a6_14[isn * nlj + jsn] * t->type_c14lj[ityp];
            *b12 = t->b12_14[isn * nlj + jsn] * t->type_c14lj[ityp];
        } else {
            *qij = t->atom_charge0[isp] * t->atom_charge0[jsp] * p->coulomb_factor;
            *a6 = t->a6_0[isn * nlj + jsn];
            *b12 = t->b12_0[isn * nlj + jsn];
        }
    } else {
        *qij = t->atom_charge[isp] * t->atom_charge[jsp] * p->coulomb_factor;
        *a6 = t->a6[isn * nlj + jsn];
        *b12 = t->b12[isn * nlj + jsn];
        if (p->ee_enabled && t->type_ee[ityp] * t->type_ee[jtyp] == 0) *rae = p->ee_raex;
    }
}

/* LOCALF (forces.f:813-998): short-list pairs into H, energies PELS2/PE2,
 * virials VIR2 and VIRFX..Z, then WIRSS over all atoms.
 * Pinned by: reference HX and energies (golden dumps), closed-form
 * two-atom LJ/Ewald cases and Newton's third law (tests).               */
void orc_localf(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, const int32_t *nbs1, const int32_t *nbs2, int64_t mbsh,
                double *hx, double *hy, double *hz, mdx_energies *e) {
    const double alphad = p->alpha;
    for (int64_t inp = 0; inp < mbsh; ++inp) {
        const int isp0 = nbs1[inp];
        const int isp = abs(isp0) - 1, jsp = nbs2[inp] - 1;
        const int ityp = t->atom_moltype[isp], jtyp = t->atom_moltype[jsp];
        const int imol = t->atom_mol[isp], jmol = t->atom_mol[jsp];
        const int mtr = t->type_pair[ityp * t->nmoltypes + jtyp];
        double qij, a6, b12, rae;
        pair_params(t, p, isp0 < 0, isp, jsp, &qij, &a6, &b12, &rae);
        double dx = x[isp] - x[jsp], dy = y[isp] - y[jsp], dz = z[isp] - z[jsp];
        min_image(&dx, &dy, &dz, p, 1);
        double rr = dx * dx + dy * dy;
        rr = rr + dz * dz;
        const double r1 = sqrt(rr);
        const double r1i = 1.0 / r1;
        const double r1e = r1 + rae;
        const double r1ei = 1.0 / r1e;
        double ee1, ees, force;
        if (isp0 <= 0 || (imol == jmol && t->type_ee[ityp] == 0)) {
            ee1 = qij * r1i;
            force = t->type_move[ityp] ? ee1 * (r1i * r1i) : 0.0;
            ees = ee1;
        } else {
            const double alphar = alphad * r1;
            const double ttt = 1.0 / (1.0 + B1 * alphar);
            const double exp2a = exp(-(alphar * alphar));
            const double erfct = ((((A5 * ttt + A4) * ttt + A3) * ttt + A2) * ttt + A1) * ttt * exp2a;
            ee1 = qij * r1ei;
            ees = ee1 * erfct;
            force = ee1 * (TOTPI * alphad * exp2a + erfct * r1ei) * r1i;
        }
        if (imol == jmol) e->pes141[ityp] += ee1;
        else e->pote1[mtr] += ee1;
        e->pels2 += ees;
        if (b12 != 0.0) {
            const double r2i = r1ei * r1ei;
            const double r6i = r2i * r2i * r2i;
            double esr = (a6 + b12 * r6i) * r6i;
            const double fsr = (6.0 * a6 + 12.0 * b12 * r6i) * r6i * r1ei;
            force = force + fsr * r1i;
            if (imol == jmol) {
                if (t->type_move[ityp]) {
                    e->psr141[ityp] += esr;
                    e->vir2 += fsr * r1;
                } else {
                    esr = 0.0;
                    force = 0.0;
                }
            } else {
                e->potl1[mtr] += esr;
                e->vir2 += fsr * r1;
            }
            e->pe2 += esr;
        }
        hx[isp] += force * dx; hy[isp] += force * dy; hz[isp] += force * dz;
        hx[jsp] -= force * dx; hy[jsp] -= force * dy; hz[jsp] -= force * dz;
        e->virfx += force * (dx * dx);
        e->virfy += force * (dy * dy);
        e->virfz += force * (dz * dz);
    }
}

/* FORCES (forces.f:1014-1202): long-list pairs into G, PELS1/PE1, VIR1,
 * VIRX..Z.  Differs from LOCALF in rounding order of the direct Coulomb
 * term (QIJ/R1), the LMOVE handling of intramolecular LJ and the guard on
 * the virial projections.  Pinned as LOCALF.                            */
void orc_forces(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, const int32_t *nbl1, const int32_t *nbl2, int64_t mbln,
                double *gx, double *gy, double *gz, mdx_energies *e) {
    const double alphad = p->alpha;
    for (int64_t inp = 0; inp < mbln; ++inp) {
        const int isp0 = nbl1[inp];
        const int isp = abs(isp0) - 1, jsp = nbl2[inp] - 1;
        const int ityp = t->atom_moltype[isp], jtyp = t->atom_moltype[jsp];
        const int imol = t->atom_mol[isp], jmol = t->atom_mol[jsp];
        const int mtr = t->type_pair[ityp * t->nmoltypes + jtyp];
        double qij, a6, b12, rae;
        pair_params(t, p, isp0 < 0, isp, jsp, &qij, &a6, &b12, &rae);
        double dx = x[isp] - x[jsp], dy = y[isp] - y[jsp], dz = z[isp] - z[jsp];
        min_image(&dx, &dy, &dz, p, 0);
        double rr = dx * dx + dy * dy;
        rr = rr + dz * dz;
        const double r1 = sqrt(rr);
        const double r1i = 1.0 / r1;
        const double r1e = r1 + rae;
        const double r1ei = 1.0 / r1e;
        double ee1, ees, force;
        if (isp0 <= 0 || (imol == jmol && t->type_ee[ityp] == 0)) {
            ee1 = qij / r1;
            force = t->type_move[ityp] ? ee1 / (r1 * r1) : 0.0;
            ees = ee1;
        } else {
            const double alphar = alphad * r1;
            const double ttt = 1.0 / (1.0 + B1 * alphar);
            const double exp2a = exp(-(alphar * alphar));
            const double erfct = ((((A5 * ttt + A4) * ttt + A3) * ttt + A2) * ttt + A1) * ttt * exp2a;
            ee1 = qij * r1ei;
            ees = ee1 * erfct;
            force = ee1 * (TOTPI * alphad * exp2a + erfct * r1ei) * r1i;
        }
        if (imol == jmol) e->pes141[ityp] += ee1;
        else e->pote1[mtr] += ee1;
        e->pels1 += ees;
        if (b12 != 0.0) {
            const double r2i = r1ei * r1ei;
            const double r6i = r2i * r2i * r2i;
            const double esr = (a6 + b12 * r6i) * r6i;
            const double fsr = (6.0 * a6 + 12.0 * b12 * r6i) * r6i * r1ei;
            force = force + fsr * r1i;
            if (imol == jmol) {
                if (t->type_move[ityp]) e->vir1 += fsr * r1;
                e->psr141[ityp] += esr;
            } else {
                e->potl1[mtr] += esr;
                e->vir1 += fsr * r1;
            }
            e->pe1 += esr;
        }
        if (imol != jmol || t->type_move[ityp]) {
            e->virx += force * (dx * dx);
            e->viry += force * (dy * dy);
            e->virz += force * (dz * dz);
        }
        gx[isp] += force * dx; gy[isp] += force * dy; gz[isp] += force * dz;
        gx[jsp] -= force * dx; gy[jsp] -= force * dy; gz[jsp] -= force * dz;
    }
}

/* ------------------------------------------------------------------ */
/* INIFUR (forces.f:1356-1449), rectangular cell: k-vectors times the box
 * edge, k_z in [-K,K], k_x in [0,K], k_y from 0 (k_z > 0) or 1 (k_z <= 0)
 * when k_x = 0, kept when |k|^2 <= 4 alpha^2 FEXP.  Returns NKV, or -1 if
 * cap is exceeded.  Self energy per molecule type into spee_self.
 * Pinned by: the reference's RKX/RKY/RKZ (golden dumps, exact) and the
 * count of lattice points in a sphere.                                  */
int64_t orc_inifur(const mdx_topology *t, const mdx_params *p, double *rkx, double *rky, double *rkz,
                   int64_t cap, double *spee_self, double *spei) {
    const double alp2 = p->alpha * p->alpha;
    const double cutk2 = 4.0 * alp2 * p->fexp;
    const double cutk = sqrt(cutk2);
    const double twopi = 2.0 * PI_R;
    const int kmax = (int)(cutk * p->box[0] / twopi + 1.0);
    const int kmay = (int)(cutk * p->box[1] / twopi + 1.0);
    const int kmaz = (int)(cutk * p->box[2] / twopi + 1.0);
    int64_t ikv = 0;
    for (int ikz = -kmaz; ikz <= kmaz; ++ikz)
        for (int ikx = 0; ikx <= kmax; ++ikx) {
            int ky1, ky2 = kmay;
            if (ikx == 0) ky1 = (ikz > 0) ? 0 : 1;
            else ky1 = -kmay;
            for (int iky = ky1; iky <= ky2; ++iky) {
                const double rxv = twopi * ikx, ryv = twopi * iky, rzv = twopi * ikz;
                const double ax = rxv / p->box[0], ay = ryv / p->box[1], az = rzv / p->box[2];
                double rk2 = ax * ax + ay * ay;
                rk2 = rk2 + az * az;
                if (rk2 <= cutk2) {
                    if (ikv < cap) { rkx[ikv] = rxv; rky[ikv] = ryv; rkz[ikv] = rzv; }
                    ikv++;
                }
            }
        }
    /* self energy (forces.f:1429-1447) */
    const double abc = p->alpha / sqrt(PI_R);
    double total = 0.0;
    for (int ityp = 0; ityp < t->nmoltypes; ++ityp) {
        const double esci = t->type_ee[ityp] == 0 ? p->ee_self_scale : 1.0;
        int nspec = 0, nsits = 0, first = -1;
        for (int i = 0; i < t->natoms; ++i)
            if (t->atom_moltype[i] == ityp) {
                if (first < 0) first = i;
                if (t->atom_mol[i] == t->atom_mol[first]) nsits++;
            }
        if (first < 0) { spee_self[ityp] = 0.0; continue; }
        int lastmol = -1;
        for (int i = 0; i < t->natoms; ++i)
            if (t->atom_moltype[i] == ityp && t->atom_mol[i] != lastmol) { nspec++; lastmol = t->atom_mol[i]; }
        double spet = 0.0;
        for (int s = 0; s < nsits; ++s) {
            const double qs = esci * t->atom_charge0[first + s];
            const double qii = p->coulomb_factor * (qs * qs);
            spet = spet - (double)nspec * abc * qii;
        }
        spee_self[ityp] = spet;
        total = total + spet;
    }
    *spei = total;
    return ikv > cap ? -1 : ikv;
}

/* FURIR (forces.f:1217-1339): reciprocal Ewald energy QPE into PELS1,
 * forces into G, virial projections.  Skipped for a single molecule
 * (forces.f:1266).  Pinned by: reference GX/PELS1 (golden dumps), the
 * rock-salt Madelung constant and alpha-independence of the total Ewald
 * energy (tests).                                                       */
void orc_furir_task(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                    const double *z, const double *rkx, const double *rky, const double *rkz, int64_t nkv,
                    int taskid, int numtask, double *gx, double *gy, double *gz, mdx_energies *e) {
    if (nkv == 0 || (t->nmoltypes == 1 && t->nmols == 1)) return;   /* forces.f:1266 */
    const int n = t->natoms;
    double *ts = (double *)malloc(sizeof(double) * n), *tc = (double *)malloc(sizeof(double) * n);
    const double alp2 = p->alpha * p->alpha;
    const double vol = p->box[0] * p->box[1] * p->box[2];
    const double cfur = 4.0 * PI_R * p->coulomb_factor / vol;
    double qpe = 0.0;
    for (int64_t k = numtask - taskid - 1; k < nkv; k += numtask) {
        const double rxv = rkx[k] / p->box[0], ryv = rky[k] / p->box[1], rzv = rkz[k] / p->box[2];
        double rk2 = rxv * rxv + ryv * ryv;
        rk2 = rk2 + rzv * rzv;
        double scs = 0.0, ssn = 0.0;
        for (int i = 0; i < n; ++i) {
            double scp = rxv * x[i] + ryv * y[i];
            scp = scp + rzv * z[i];
            const double qq = t->atom_charge[i];
            const double sinsc = qq * sin(scp), cossc = qq * cos(scp);
            ts[i] = sinsc; tc[i] = cossc;
            ssn = ssn + sinsc; scs = scs + cossc;
        }
        const double akc = cfur * exp(-0.25 * rk2 / alp2) / rk2;
        const double s2 = ssn * ssn + scs * scs;
        qpe = qpe + akc * s2;
        const double ak2 = 2.0 * akc;
        for (int i = 0; i < n; ++i) {
            const double qfor = ak2 * (scs * ts[i] - ssn * tc[i]);
            gx[i] += rxv * qfor; gy[i] += ryv * qfor; gz[i] += rzv * qfor;
        }
        const double akv = 0.5 * akc * (4.0 * alp2 + rk2) / (alp2 * rk2);
        e->virx = e->virx - rxv * rxv * akv * s2;
        e->viry = e->viry - ryv * ryv * akv * s2;
        e->virz = e->virz - rzv * rzv * akv * s2;
    }
    e->virx += qpe; e->viry += qpe; e->virz += qpe;
    e->pels1 += qpe;
    free(ts); free(tc);
}

void orc_furir(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
               const double *z, const double *rkx, const double *rky, const double *rkz, int64_t nkv,
               double *gx, double *gy, double *gz, mdx_energies *e) {
    orc_furir_task(t, p, x, y, z, rkx, rky, rkz, nkv, 0, 1, gx, gy, gz, e);
}

/* One exclusion correction pair of ETERMS (forces.f:1554-1577). */
static void eterm_pair(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                       const double *z, int isp, int jsp, int ityp, double *gx, double *gy, double *gz,
                       mdx_energies *e) {
    const double qij = p->coulomb_factor * t->atom_charge[isp] * t->atom_charge[jsp];
    double dx = x[isp] - x[jsp], dy = y[isp] - y[jsp], dz = z[isp] - z[jsp];
    min_image(&dx, &dy, &dz, p, 0);
    double r2 = dx * dx + dy * dy;
    r2 = r2 + dz * dz;
    const double r2i = 1.0 / r2;
    const double r1 = sqrt(r2);
    const double r1i = r1 * r2i;
    const double alphar = p->alpha * r1;
    const double ttt = 1.0 / (1.0 + B1 * alphar);
    const double exp2a = exp(-(alphar * alphar));
    const double erfc_ = ((((A5 * ttt + A4) * ttt + A3) * ttt + A2) * ttt + A1) * ttt * exp2a;
    const double ees = qij * (erfc_ - 1.0) * r1i;
    e->spe += ees;
    e->spee[ityp] += ees;
    if (t->type_move[ityp]) {
        const double fes = qij * ((TOTPI * alphar * exp2a + erfc_) - 1.0) * r1i * r2i;
        gx[isp] += fes * dx; gy[isp] += fes * dy; gz[isp] += fes * dz;
        gx[jsp] -= fes * dx; gy[jsp] -= fes * dy; gz[jsp] -= fes * dz;
        e->virx += fes * (dx * dx);
        e->viry += fes * (dy * dy);
        e->virz += fes * (dz * dz);
    }
}

/* ETERMS (forces.f:1482-1676): SPE = SPEI plus, for every bound pair of a
 * normal molecule (1-2, 1-3, 1-4; larger index first), the removal of the
 * reciprocal-space interaction; for lambda-coupled types every
 * intramolecular pair.  Pinned by: reference SPE/GX (golden dumps) and
 * the closed form q_i q_j (erfc(ar) - 1)/r.                            */
void orc_eterms(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, const double *spee_self, double spei, double *gx, double *gy,
                double *gz, mdx_energies *e) {
    for (int i = 0; i < t->nmoltypes; ++i) e->spee[i] = spee_self[i];
    e->spe = spei;
    for (int isp = 0; isp < t->natoms; ++isp) {
        const int ityp = t->atom_moltype[isp];
        if (t->type_ee[ityp] == 0) continue;
        for (int k = t->excl_start[isp]; k < t->excl_start[isp + 1]; ++k) {
            const int jsp = t->excl_atom[k];
            if (isp > jsp) eterm_pair(t, p, x, y, z, isp, jsp, ityp, gx, gy, gz, e);
        }
    }
    for (int ityp = 0; ityp < t->nmoltypes; ++ityp) {
        if (t->type_ee[ityp] != 0) continue;
        for (int a = 0; a < t->natoms;) {
            if (t->atom_moltype[a] != ityp) { ++a; continue; }
            int b = a;
            while (b < t->natoms && t->atom_mol[b] == t->atom_mol[a]) ++b;
            for (int ism = 1; ism < b - a; ++ism)
                for (int jsm = 0; jsm < ism; ++jsm)
                    eterm_pair(t, p, x, y, z, a + ism, a + jsm, ityp, gx, gy, gz, e);
            a = b;
        }
    }
}

/* ------------------------------------------------------------------ */
/* GETBND (forces.f:113-204).  Pinned by: reference HX/PINT/VIRB (golden
 * dumps), finite differences of the energy and sum(F) = 0 (tests).      */
void orc_getbnd(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, double *hx, double *hy, double *hz, mdx_energies *e) {
    for (int n = 0; n < t->nbonds; ++n) {
        const int i = t->bond_i[n], j = t->bond_j[n], m = t->bond_param[n];
        const int ityp = t->atom_moltype[i];
        const double req = t->bp_r0[m];
        double bx = x[j] - x[i], by = y[j] - y[i], bz = z[j] - z[i];
        pbc(&bx, &by, &bz, p);
        double bsq = bx * bx + by * by;
        bsq = bsq + bz * bz;
        const double bd = sqrt(bsq);
        const double r1i = 1.0 / bd;
        const double bbb = bd - req;
        double ebdd, forb;
        if (!t->bp_morse[m] || bbb > req) {
            const double dbnd = bbb * t->bp_k[m];
            ebdd = dbnd * bbb;
            forb = -2.0 * dbnd * r1i;
        } else {
            const double ddis = t->bp_d[m], roh = t->bp_rho[m];
            const double exm = exp(-roh * bbb);
            const double exmp = 1.0 - exm;
            ebdd = ddis * (exmp * exmp);
            forb = -2.0 * ddis * roh * exm * exmp * r1i;
        }
        e->pint += ebdd;
        if (t->type_move[ityp]) {
            hx[i] -= bx * forb; hy[i] -= by * forb; hz[i] -= bz * forb;
            hx[j] += bx * forb; hy[j] += by * forb; hz[j] += bz * forb;
            e->virb += forb * bsq;
            e->virfx += forb * (bx * bx);
            e->virfy += forb * (by * by);
            e->virfz += forb * (bz * bz);
        }
    }
}

/* STRBND (forces.f:530-800): flexible SPC water, Morse (IPOT 2) or
 * harmonic (IPOT 1) O-H terms, harmonic H-H and fixed cross terms, per
 * molecule.  The Morse energy is kept beyond R_OH > BBMA (a REAL*4 1.3)
 * while the force switches to the harmonic form -2 s D rho1^2 -- for
 * both bonds with rho1, as the reference writes it; the harmonic variant
 * takes no periodic image.  Pinned by: the reference's HX/PINT/VIRB of an
 * fSPC box (golden dump fspc_tiny, IPOT 2, bonds off equilibrium) and, for
 * both variants, forces equal to minus the central difference of PINT
 * away from the switch (tests/test_oracle_math.py).  IPOT 1 cannot be set
 * by an input (input.f:1304-1306): no reference output exists for it, so
 * the gradient check is its only pin.                                   */
void orc_strbnd(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, double *hx, double *hy, double *hz, mdx_energies *e) {
    const double bbm = (double)1.3f;
    for (int n = 0; n < t->nspc; ++n) {
        const int i = t->spc_atom[3 * n], j = t->spc_atom[3 * n + 1], k = t->spc_atom[3 * n + 2];
        const double *q = t->spc_par + 12 * (size_t)n;
        const double roh1 = q[0], roh2 = q[1], rhh = q[2], aaa = q[3], bbb = q[4], ccc = q[5];
        const double rho1 = q[6], rho2 = q[7], eee = q[8], ggg = q[9];
        const int ipt = (int)q[10], lmove = q[11] != 0.0;
        const double fff = eee;
        double ax = x[j] - x[i], ay = y[j] - y[i], az = z[j] - z[i];
        double bx = x[k] - x[i], by = y[k] - y[i], bz = z[k] - z[i];
        double cx = x[k] - x[j], cy = y[k] - y[j], cz = z[k] - z[j];
        if (ipt == 2) { pbc(&ax, &ay, &az, p); pbc(&bx, &by, &bz, p); pbc(&cx, &cy, &cz, p); }
        const double rra = ax * ax + ay * ay + az * az;
        const double rrb = bx * bx + by * by + bz * bz;
        const double rrc = cx * cx + cy * cy + cz * cz;
        const double ass = sqrt(rra), bss = sqrt(rrb), css = sqrt(rrc);
        const double sa = ass - roh1, sb = bss - roh2, sh = css - rhh;
        double wb1, wb2, fa, fb;
        if (ipt == 2) {
            const double expa = exp(-rho1 * sa), expam = 1.0 - expa;
            const double expb = exp(-rho2 * sb), expbm = 1.0 - expb;
            wb1 = aaa * (expam * expam) + eee * sa * sh;
            wb2 = bbb * (expbm * expbm) + fff * sb * sh;
            const double aadd = aaa * rho1, bbdd = bbb * rho2;
            fa = ass > bbm ? -2. * sa * aadd * rho1 : -2. * aadd * expa * expam;
            fa = fa - eee * sh - ggg * sb;
            fb = bss > bbm ? -2. * sb * bbdd * rho1 : -2. * bbdd * expb * expbm;
            fb = fb - fff * sh - ggg * sa;
        } else {
            wb1 = aaa * sa * sa + eee * sa * sh;
            wb2 = bbb * sb * sb + fff * sb * sh;
            fa = -2.0 * aaa * sa - eee * sh - ggg * sb;
            fb = -2.0 * bbb * sb - fff * sh - ggg * sa;
        }
        const double wb3 = ccc * sh * sh + ggg * sa * sb;
        e->pint = e->pint + wb1 + wb2 + wb3;
        if (!lmove) continue;
        double fc = -2.0 * ccc * sh - eee * sa - fff * sb;
        fa = fa / ass; fb = fb / bss; fc = fc / css;
        const double axx = fa * ax, ayy = fa * ay, azz = fa * az;
        const double bxx = fb * bx, byy = fb * by, bzz = fb * bz;
        const double cxx = fc * cx, cyy = fc * cy, czz = fc * cz;
        hx[i] = hx[i] - axx - bxx; hy[i] = hy[i] - ayy - byy; hz[i] = hz[i] - azz - bzz;
        hx[j] = hx[j] + axx - cxx; hy[j] = hy[j] + ayy - cyy; hz[j] = hz[j] + azz - czz;
        hx[k] = hx[k] + bxx + cxx; hy[k] = hy[k] + byy + cyy; hz[k] = hz[k] + bzz + czz;
        const double vix = ax * axx + bx * bxx + cx * cxx;
        const double viy = ay * ayy + by * byy + cy * cyy;
        const double viz = az * azz + bz * bzz + cz * czz;
        e->virb = e->virb + vix + viy + viz;
        e->virfx += vix; e->virfy += viy; e->virfz += viz;
    }
}

/* GETANG (forces.f:216-309).  Pinned as GETBND.                         */
void orc_getang(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, double *hx, double *hy, double *hz, mdx_energies *e) {
    for (int n = 0; n < t->nangles; ++n) {
        const int i = t->angle_i[n], j = t->angle_j[n], k = t->angle_k[n], m = t->angle_param[n];
        const int ityp = t->atom_moltype[i];
        const double fcn = t->ap_k[m], aeq = t->ap_theta0[m];
        double ax = x[i] - x[j], ay = y[i] - y[j], az = z[i] - z[j];
        pbc(&ax, &ay, &az, p);
        double asq = ax * ax + ay * ay; asq = asq + az * az;
        const double asqi = 1.0 / asq;
        double bx = x[k] - x[j], by = y[k] - y[j], bz = z[k] - z[j];
        pbc(&bx, &by, &bz, p);
        double bsq = bx * bx + by * by; bsq = bsq + bz * bz;
        const double bsqi = 1.0 / bsq;
        const double ab = sqrt(asq * bsq);
        const double abi = 1.0 / ab;
        double cosa = ax * bx + ay * by; cosa = cosa + az * bz;
        cosa = cosa * abi;
        if (cosa > 1.0) cosa = 1.0;
        if (cosa < -1.0) cosa = -1.0;
        const double alf = acos(cosa);
        const double sina = sin(alf);
        const double sinai = 1.0 / sina;
        const double dalf = alf - aeq;
        const double abc = dalf * fcn;
        const double eang = abc * dalf;
        e->pint += eang;
        if (t->type_move[ityp]) {
            const double sab = -2.0 * abc * sinai;
            const double cab = sab * cosa;
            const double fab = sab * abi;
            const double faa = cab * asqi, fbb = cab * bsqi;
            const double fax = fab * bx - faa * ax, fay = fab * by - faa * ay, faz = fab * bz - faa * az;
            const double fbx = fab * ax - fbb * bx, fby = fab * ay - fbb * by, fbz = fab * az - fbb * bz;
            hx[i] -= fax; hy[i] -= fay; hz[i] -= faz;
            hx[j] += fax + fbx; hy[j] += fay + fby; hz[j] += faz + fbz;
            hx[k] -= fbx; hy[k] -= fby; hz[k] -= fbz;
            e->virfx = e->virfx - ax * fax - bx * fbx;
            e->virfy = e->virfy - ay * fay - by * fby;
            e->virfz = e->virfz - az * faz - bz * fbz;
        }
    }
}

/* GETTRS (forces.f:323-513).  Pinned by: reference HX/PINT (golden dumps
 * of the chain system), finite differences, and the dihedral of known
 * cis/trans geometries (tests).                                         */
void orc_gettrs(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, double *hx, double *hy, double *hz, mdx_energies *e) {
    for (int n = 0; n < t->ntors; ++n) {
        const int i = t->tors_i[n], j = t->tors_j[n], k = t->tors_k[n], l = t->tors_l[n];
        const int m = t->tors_param[n];
        const int ityp = t->atom_moltype[i];
        double ax = x[j] - x[i], ay = y[j] - y[i], az = z[j] - z[i];
        double bx = x[k] - x[j], by = y[k] - y[j], bz = z[k] - z[j];
        double cx = x[l] - x[k], cy = y[l] - y[k], cz = z[l] - z[k];
        pbc(&ax, &ay, &az, p); pbc(&bx, &by, &bz, p); pbc(&cx, &cy, &cz, p);
        double ab = ax * bx + ay * by; ab = ab + az * bz;
        double bc = bx * cx + by * cy; bc = bc + bz * cz;
        double ac = ax * cx + ay * cy; ac = ac + az * cz;
        double at = ax * ax + ay * ay; at = at + az * az;
        double bt = bx * bx + by * by; bt = bt + bz * bz;
        double ct = cx * cx + cy * cy; ct = ct + cz * cz;
        double axb = (at * bt) - (ab * ab);
        double bxc = (bt * ct) - (bc * bc);
        const double fnum = (ab * bc) - (ac * bt);
        double den = axb * bxc;
        if (!(den > 1.0e-10)) continue;
        den = sqrt(den);
        double co = fnum / den;
        if (co > 1.0) co = 1.0;
        if (co < -1.0) co = -1.0;
        double signum = ax * (by * cz - cy * bz) + ay * (bz * cx - cz * bx);
        signum = signum + az * (bx * cy - cx * by);
        const double arg = fsign(acos(co), signum);
        const double sin1f = sin(arg);
        double si = sin1f;
        if (fabs(si) < 1.0e-10) si = fsign(1.0e-10, si);
        double pott, deri;
        switch (t->tp_kind[m]) {
        case 0: {
            const double earg = t->tp_mult[m] * arg - t->tp_delta[m];
            pott = t->tp_k[m] * (1.0 + cos(earg));
            deri = -t->tp_k[m] * t->tp_mult[m] * sin(earg);
            break;
        }
        case 1: {
            const double fn1 = t->tp_k1[m], fn2 = t->tp_k2[m], fn3 = t->tp_k3[m];
            const double cos1f = cos(arg);
            const double cos2f = 2.0 * (cos1f * cos1f) - 1.0;
            const double cos3f = cos1f * (2.0 * cos2f - 1.0);
            const double sin2f = 2.0 * sin1f * cos1f;
            const double sin3f = sin1f * cos2f + sin2f * cos1f;
            pott = fn1 * (1.0 + cos1f) + fn2 * (1.0 - cos2f) + fn3 * (1.0 + cos3f);
            deri = -fn1 * sin1f + 2.0 * fn2 * sin2f - 3.0 * fn3 * sin3f;
            break;
        }
        case 5: {
            const double earg = arg - t->tp_delta[m];
            const double c1 = cos(earg), c2 = c1 * c1, c3 = c1 * c2, c4 = c1 * c3, c5 = c1 * c4;
            const double s1 = sin(earg);
            pott = t->tp_k[m] + t->tp_k1[m] * c1 + t->tp_k2[m] * c2 + t->tp_k3[m] * c3 +
                   t->tp_k4[m] * c4 + t->tp_k5[m] * c5;
            deri = -s1 * (t->tp_k1[m] + 2.0 * t->tp_k2[m] * c1 + 3.0 * t->tp_k3[m] * c2 +
                          4.0 * t->tp_k4[m] * c3 + 5.0 * t->tp_k5[m] * c4);
            break;
        }
        case -1: {
            const double earg = arg - t->tp_delta[m];
            pott = t->tp_k[m] * (earg * earg);
            deri = 2.0 * t->tp_k[m] * earg;
            break;
        }
        default:
            continue;
        }
        e->pint += pott;
        if (!t->type_move[ityp]) continue;
        const double de1 = deri / den / si;
        axb = axb / den * co;
        bxc = bxc / den * co;
        double dnum, dden;
        /* x */
        dnum = cx * bt - bx * bc;
        dden = (ab * bx - ax * bt) * bxc;
        const double ffi1 = (dnum - dden) * de1;
        dnum = ((bx - ax) * bc - ab * cx) + (2.0 * ac * bx - cx * bt);
        dden = axb * (bc * cx - bx * ct) + (ax * bt - at * bx - ab * (bx - ax)) * bxc;
        const double ffj1 = (dnum - dden) * de1;
        dnum = ab * bx - ax * bt;
        dden = axb * (bt * cx - bc * bx);
        const double ffl1 = (dnum - dden) * de1;
        const double ffk1 = -(ffi1 + ffj1 + ffl1);
        /* y */
        dnum = cy * bt - by * bc;
        dden = (ab * by - ay * bt) * bxc;
        const double ffi2 = (dnum - dden) * de1;
        dnum = ((by - ay) * bc - ab * cy) + (2.0 * ac * by - cy * bt);
        dden = axb * (bc * cy - by * ct) + (ay * bt - at * by - ab * (by - ay)) * bxc;
        const double ffj2 = (dnum - dden) * de1;
        dnum = ab * by - ay * bt;
        dden = axb * (bt * cy - bc * by);
        const double ffl2 = (dnum - dden) * de1;
        const double ffk2 = -(ffi2 + ffj2 + ffl2);
        /* z */
        dnum = cz * bt - bz * bc;
        dden = (ab * bz - az * bt) * bxc;
        const double ffi3 = (dnum - dden) * de1;
        dnum = ((bz - az) * bc - ab * cz) + (2.0 * ac * bz - cz * bt);
        dden = axb * (bc * cz - bz * ct) + (az * bt - at * bz - ab * (bz - az)) * bxc;
        const double ffj3 = (dnum - dden) * de1;
        dnum = ab * bz - az * bt;
        dden = axb * (bt * cz - bc * bz);
        const double ffl3 = (dnum - dden) * de1;
        const double ffk3 = -(ffi3 + ffj3 + ffl3);
        hx[i] += ffi1; hx[j] += ffj1; hx[k] += ffk1; hx[l] += ffl1;
        hy[i] += ffi2; hy[j] += ffj2; hy[k] += ffk2; hy[l] += ffl2;
        hz[i] += ffi3; hz[j] += ffj3; hz[k] += ffk3; hz[l] += ffl3;
        e->virfx = e->virfx - (ax + bx) * ffi1 - bx * ffj1 + cx * ffl1;
        e->virfy = e->virfy - (ay + by) * ffi2 - by * ffj2 + cy * ffl2;
        e->virfz = e->virfz - (az + bz) * ffi3 - bz * ffj3 + cz * ffl3;
    }
}

/* ------------------------------------------------------------------ */
/* Molecular virial sum over atoms, -sum F.(s - X_mol) (forces.f:991-995,
 * scalar.f:72-76).                                                      */
static double mol_virial(const mdx_topology *t, const double *x, const double *y, const double *z,
                         const double *fx, const double *fy, const double *fz, const double *cx,
                         const double *cy, const double *cz) {
    double w = 0.0;
    for (int n = 0; n < t->natoms; ++n) {
        const int m = t->atom_mol[n];
        w = w - fx[n] * (x[n] - cx[m]) - fy[n] * (y[n] - cy[m]) - fz[n] * (z[n] - cz[m]);
    }
    return w;
}

/* ALLFORCE / FFORCES / SLFORCES (forces.f:36-100) on given lists.
 * group: MDX_FORCES_ALL/FAST/SLOW.  g and h are zeroed here (ZEROFS);
 * f = g + h for ALL (CM_ADDLA).  cx/cy/cz: molecular centres of mass.   */
int orc_compute(const mdx_topology *t, const mdx_params *p, const double *x, const double *y,
                const double *z, const double *cx, const double *cy, const double *cz,
                const int32_t *nbs1, const int32_t *nbs2, int64_t mbsh, const int32_t *nbl1,
                const int32_t *nbl2, int64_t mbln, const double *rkx, const double *rky,
                const double *rkz, int64_t nkv, const double *spee_self, double spei, int32_t group,
                double *fx, double *fy, double *fz, double *gx, double *gy, double *gz, double *hx,
                double *hy, double *hz, mdx_energies *e) {
    const int n = t->natoms;
    memset(e, 0, sizeof(*e));
    if (group == MDX_FORCES_ALL || group == MDX_FORCES_FAST) {
        memset(hx, 0, sizeof(double) * n); memset(hy, 0, sizeof(double) * n); memset(hz, 0, sizeof(double) * n);
        orc_localf(t, p, x, y, z, nbs1, nbs2, mbsh, hx, hy, hz, e);
        e->wirss = mol_virial(t, x, y, z, hx, hy, hz, cx, cy, cz);
        orc_strbnd(t, p, x, y, z, hx, hy, hz, e);   /* GETBND calls STRBND first */
        orc_getbnd(t, p, x, y, z, hx, hy, hz, e);
        orc_getang(t, p, x, y, z, hx, hy, hz, e);
        orc_gettrs(t, p, x, y, z, hx, hy, hz, e);
        e->npairs_short = mbsh;
    }
    if (group == MDX_FORCES_ALL || group == MDX_FORCES_SLOW) {
        memset(gx, 0, sizeof(double) * n); memset(gy, 0, sizeof(double) * n); memset(gz, 0, sizeof(double) * n);
        if (p->alpha > 0.0) {
            orc_furir(t, p, x, y, z, rkx, rky, rkz, nkv, gx, gy, gz, e);
            orc_eterms(t, p, x, y, z, spee_self, spei, gx, gy, gz, e);
        }
        orc_forces(t, p, x, y, z, nbl1, nbl2, mbln, gx, gy, gz, e);
        e->wirs = mol_virial(t, x, y, z, gx, gy, gz, cx, cy, cz);
        e->npairs_long = mbln;
    }
    if (fx) {
        for (int i = 0; i < n; ++i) {
            if (group == MDX_FORCES_ALL) { fx[i] = gx[i] + hx[i]; fy[i] = gy[i] + hy[i]; fz[i] = gz[i] + hz[i]; }
            else if (group == MDX_FORCES_FAST) { fx[i] = hx[i]; fy[i] = hy[i]; fz[i] = hz[i]; }
            else { fx[i] = gx[i]; fy[i] = gy[i]; fz[i] = gz[i]; }
        }
    }
    return MDX_OK;
}
