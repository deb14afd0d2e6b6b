// Bonded terms, Ewald exclusion corrections, force assembly and reductions.
//
// k_eterms   ETERMS  forces.f:1482-1676 (one bound/intramolecular pair per thread)
// k_bonds    GETBND  forces.f:113-204
// k_strbnd   STRBND  forces.f:530-800 (fSPC water, one molecule per thread)
// k_angles   GETANG  forces.f:216-309
// k_tors     GETTRS  forces.f:323-513
// All in FP64 (the bonded forms derived in each body); forces are scattered
// in 64-bit fixed point (order independent), energies reduced per block in a
// fixed tree order.
// k_combine  F = G + H (CM_ADDLA, scalar.f:82-116) and the molecular virial
//            sums WIRS/WIRSS (scalar.f:72-76, forces.f:991-995)
// k_springs  path-integral bead springs (PIFORCES, pimd.f:321-394)
// k_ereduce, k_ereduce_last  final FP64 sum of all per-block partials
#include "mdx_internal.h"

namespace mdx {

// Block sums of NQ values in a fixed order: a butterfly inside each
// wavefront (xor 32, 16, ..., 1), then the wavefronts' sums in wavefront
// order (blocks of at most 1024 threads); one barrier for all NQ values.
template <int NQ>
__device__ inline void block_reduce_store(double (&q)[NQ], const int (&slot)[NQ], double *out) {
    __shared__ double wred[16][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = 0; i < NQ; ++i) {
        double v = q[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) wred[w][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int nw = (blockDim.x + 63) >> 6;
        double t = wred[0][threadIdx.x];
        for (int k = 1; k < nw; ++k) t += wred[k][threadIdx.x];
        out[slot[threadIdx.x]] = t;
    }
    __syncthreads();
}

__device__ inline void zero_slots(double *o) {
    if (threadIdx.x == 0)
        for (int k = 0; k < Q_N; ++k) o[k] = 0.0;
    __syncthreads();
}


// pairs: (i, j) with i > j
__device__ inline void eterms_body(const EtermArgs &t, int blk) {
    const ScatterArgs &a = t.s;
    const int k = blk * blockDim.x + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    const double alpha = t.alpha;
    if (k < t.np) {
        const int isp = t.pairs[k].x, jsp = t.pairs[k].y;
        const double qij = t.coulf * t.q[isp] * t.q[jsp];
        double dx = a.x[isp] - a.x[jsp], dy = a.y[isp] - a.y[jsp], dz = a.z[isp] - a.z[jsp];
        min_image(dx, dy, dz, a.box);
        double r2 = dx * dx + dy * dy;
        r2 = r2 + dz * dz;
        const double r2i = 1.0 / r2;
        const double r1 = sqrt(r2);
        const double r1i = r1 * r2i;
        const double alphar = alpha * r1;
        const double ttt = 1.0 / (1.0 + kB1 * alphar);
        const double exp2a = exp(-(alphar * alphar));
        const double erfc_ = ((((kA5 * ttt + kA4) * ttt + kA3) * ttt + kA2) * ttt + kA1) * ttt * exp2a;
        v[0] = qij * (erfc_ - 1.0) * r1i;
        const double fes = qij * ((kTOTPI * alphar * exp2a + erfc_) - 1.0) * r1i * r2i;
        fx_add(a.fx + isp, fes * dx, a.fscale); fx_add(a.fy + isp, fes * dy, a.fscale);
        fx_add(a.fz + isp, fes * dz, a.fscale);
        fx_add(a.fx + jsp, -fes * dx, a.fscale); fx_add(a.fy + jsp, -fes * dy, a.fscale);
        fx_add(a.fz + jsp, -fes * dz, a.fscale);
        v[1] = fes * (dx * dx); v[2] = fes * (dy * dy); v[3] = fes * (dz * dz);
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[4] = {Q_SPE, Q_VIRX, Q_VIRY, Q_VIRZ};
    block_reduce_store<4>(v, slots, o);
}


// Bond stretch of GETBND (forces.f:113-204): with d = r_j - r_i (minimum
// image) and u = |d| - r0, E = K u^2, or the Morse well D (1 - e^{-rho u})^2
// when the bond type is Morse and u <= r0 (the harmonic form beyond).  The
// pair force follows from dE/dr: F_j = -dE/dr d/|d| = -F_i.  Virials: VIRB
// gets d . F_j = -|d| dE/dr, the projections d_x F_jx etc.
__device__ inline void bonds_body(const BondArgs &t, int blk) {
    const ScatterArgs &a = t.s;
    const int n = blk * blockDim.x + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};
    if (n < t.n) {
        const int i = t.bi[n], j = t.bj[n], m = t.bpar[n];
        double d[3] = {a.x[j] - a.x[i], a.y[j] - a.y[i], a.z[j] - a.z[i]};
        min_image(d[0], d[1], d[2], a.box);
        double len2 = d[0] * d[0] + d[1] * d[1];
        len2 = len2 + d[2] * d[2];
        const double len = sqrt(len2);
        const double u = len - t.r0[m];
        double energy, slope;                  // E and dE/dr
        if (t.morse[m] && u <= t.r0[m]) {
            const double w = exp(-t.rho[m] * u);
            energy = t.d[m] * (1.0 - w) * (1.0 - w);
            slope = 2.0 * t.d[m] * t.rho[m] * w * (1.0 - w);
        } else {
            energy = t.k[m] * u * u;
            slope = 2.0 * t.k[m] * u;
        }
        const double g = slope / len;          // F_j = -g d, F_i = +g d
        double fj[3];
        for (int c = 0; c < 3; ++c) fj[c] = -g * d[c];
        v[0] = energy;
        t.tv[n] = make_double2(len, energy);
        v[1] = -slope * len;
        for (int c = 0; c < 3; ++c) v[2 + c] = d[c] * fj[c];
        unsigned long long *F[3] = {a.fx, a.fy, a.fz};
        for (int c = 0; c < 3; ++c) {
            fx_add(F[c] + i, -fj[c], a.fscale);
            fx_add(F[c] + j, fj[c], a.fscale);
        }
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[5] = {Q_PINT, Q_VIRB, Q_VIRFX, Q_VIRFY, Q_VIRFZ};
    block_reduce_store<5>(v, slots, o);
}


// Morse (IPOT 2, minimum image) or harmonic (IPOT 1, no image) O-H
// stretches, harmonic H-H and the cross terms E sa sh, E sb sh, G sa sb.
// Beyond R_OH > 1.3 A (BBMA, a REAL*4 constant) the Morse force takes the
// harmonic form -2 s A rho1^2 while the energy stays Morse; the reference
// uses rho1 for both bonds there, and so does this kernel.
__device__ inline void strbnd_body(const SpcArgs &t, int blk) {
    const ScatterArgs &a = t.s;
    const int n = blk * blockDim.x + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};
    if (n < t.n) {
        const int i = t.atom[3 * n], j = t.atom[3 * n + 1], k = t.atom[3 * n + 2];
        const double *q = t.par + 12 * (size_t)n;
        const double roh1 = q[0], roh2 = q[1], rhh = q[2], aaa = q[3], bbb = q[4], ccc = q[5];
        const double rho1 = q[6], rho2 = q[7], eee = q[8], ggg = q[9];
        const bool morse = q[10] == 2.0, lmove = q[11] != 0.0;
        double ax = a.x[j] - a.x[i], ay = a.y[j] - a.y[i], az = a.z[j] - a.z[i];
        double bx = a.x[k] - a.x[i], by = a.y[k] - a.y[i], bz = a.z[k] - a.z[i];
        double cx = a.x[k] - a.x[j], cy = a.y[k] - a.y[j], cz = a.z[k] - a.z[j];
        if (morse) {
            min_image(ax, ay, az, a.box); min_image(bx, by, bz, a.box); min_image(cx, cy, cz, a.box);
        }
        const double ass = sqrt(ax * ax + ay * ay + az * az);
        const double bss = sqrt(bx * bx + by * by + bz * bz);
        const double css = sqrt(cx * cx + cy * cy + cz * cz);
        const double sa = ass - roh1, sb = bss - roh2, sh = css - rhh;
        double wa, wb, fa, fb;
        if (morse) {
            const double bbm = (double)1.3f;
            const double ea = exp(-rho1 * sa), eam = 1.0 - ea;
            const double eb = exp(-rho2 * sb), ebm = 1.0 - eb;
            wa = aaa * (eam * eam) + eee * sa * sh;
            wb = bbb * (ebm * ebm) + eee * sb * sh;
            const double ad = aaa * rho1, bd = bbb * rho2;
            fa = ass > bbm ? -2. * sa * ad * rho1 : -2. * ad * ea * eam;
            fb = bss > bbm ? -2. * sb * bd * rho1 : -2. * bd * eb * ebm;
        } else {
            wa = aaa * sa * sa + eee * sa * sh;
            wb = bbb * sb * sb + eee * sb * sh;
            fa = -2.0 * aaa * sa;
            fb = -2.0 * bbb * sb;
        }
        fa = (fa - eee * sh - ggg * sb) / ass;
        fb = (fb - eee * sh - ggg * sa) / bss;
        const double fc = (-2.0 * ccc * sh - eee * sa - eee * sb) / css;
        const double wc = ccc * sh * sh + ggg * sa * sb;
        v[0] = wa + wb + wc;
        t.tv[3 * n] = make_double2(ass, wa);
        t.tv[3 * n + 1] = make_double2(bss, wb);
        t.tv[3 * n + 2] = make_double2(css, wc);
        if (lmove) {
            const double fax = fa * ax, fay = fa * ay, faz = fa * az;
            const double fbx = fb * bx, fby = fb * by, fbz = fb * bz;
            const double fcx = fc * cx, fcy = fc * cy, fcz = fc * cz;
            fx_add(a.fx + i, -fax - fbx, a.fscale); fx_add(a.fy + i, -fay - fby, a.fscale);
            fx_add(a.fz + i, -faz - fbz, a.fscale);
            fx_add(a.fx + j, fax - fcx, a.fscale); fx_add(a.fy + j, fay - fcy, a.fscale);
            fx_add(a.fz + j, faz - fcz, a.fscale);
            fx_add(a.fx + k, fbx + fcx, a.fscale); fx_add(a.fy + k, fby + fcy, a.fscale);
            fx_add(a.fz + k, fbz + fcz, a.fscale);
            v[2] = ax * fax + bx * fbx + cx * fcx;
            v[3] = ay * fay + by * fby + cy * fcy;
            v[4] = az * faz + bz * fbz + cz * fcz;
            v[1] = v[2] + v[3] + v[4];
        }
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[5] = {Q_PINT, Q_VIRB, Q_VIRFX, Q_VIRFY, Q_VIRFZ};
    block_reduce_store<5>(v, slots, o);
}


// Bending of GETANG (forces.f:216-309): E = K (theta - theta0)^2 with the
// angle theta at j between p = r_i - r_j and s = r_k - r_j (minimum images).
// With c = cos theta = p.s / (|p| |s|), d c / d p = (s/|s| - c p/|p|) / |p|
// and d theta = -d c / sin theta, so
//   F_i = (dE/dtheta / sin theta) (s/|s| - c p/|p|) / |p|,
//   F_k = (dE/dtheta / sin theta) (p/|p| - c s/|s|) / |s|,  F_j = -F_i - F_k.
// c is clamped to [-1, 1] as the reference does; sin theta is sqrt(1 - c^2).
// Virial projections: p_x F_ix + s_x F_kx (positions relative to j).
__device__ inline void angles_body(const AngleArgs &t, int blk) {
    const ScatterArgs &a = t.s;
    const int n = blk * blockDim.x + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (n < t.n) {
        const int i = t.ai[n], j = t.aj[n], k = t.ak[n], m = t.apar[n];
        double p[3] = {a.x[i] - a.x[j], a.y[i] - a.y[j], a.z[i] - a.z[j]};
        double s[3] = {a.x[k] - a.x[j], a.y[k] - a.y[j], a.z[k] - a.z[j]};
        min_image(p[0], p[1], p[2], a.box);
        min_image(s[0], s[1], s[2], a.box);
        double pp = p[0] * p[0] + p[1] * p[1];
        pp = pp + p[2] * p[2];
        double ss = s[0] * s[0] + s[1] * s[1];
        ss = ss + s[2] * s[2];
        double ps = p[0] * s[0] + p[1] * s[1];
        ps = ps + p[2] * s[2];
        const double lp = sqrt(pp), ls = sqrt(ss);
        const double c = fmax(-1.0, fmin(1.0, ps / (lp * ls)));
        const double theta = acos(c);
        const double dev = theta - t.theta0[m];
        v[0] = t.k[m] * dev * dev;
        t.tv[n] = make_double2(theta, v[0]);
        const double h = 2.0 * t.k[m] * dev / sqrt(1.0 - c * c);   // dE/dtheta / sin theta
        double fi[3], fk[3];
        for (int q = 0; q < 3; ++q) {
            const double ph = p[q] / lp, sh = s[q] / ls;
            fi[q] = h * (sh - c * ph) / lp;
            fk[q] = h * (ph - c * sh) / ls;
        }
        unsigned long long *F[3] = {a.fx, a.fy, a.fz};
        for (int q = 0; q < 3; ++q) {
            fx_add(F[q] + i, fi[q], a.fscale);
            fx_add(F[q] + j, -(fi[q] + fk[q]), a.fscale);
            fx_add(F[q] + k, fk[q], a.fscale);
            v[1 + q] = p[q] * fi[q] + s[q] * fk[q];
        }
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[4] = {Q_PINT, Q_VIRFX, Q_VIRFY, Q_VIRFZ};
    block_reduce_store<4>(v, slots, o);
}


// Dihedral terms of GETTRS (forces.f:323-513).  Bond vectors b1 = r_j - r_i,
// b2 = r_k - r_j, b3 = r_l - r_k (minimum images), plane normals m = b1 x b2,
// nn = b2 x b3; the IUPAC angle (trans = 180 deg, sign of b1 . nn) is
// phi = atan2(|b2| b1 . nn, m . nn).  Energies of phi by the type's form
// (ITORS 0: K (1 + cos(M phi - delta)); 1: K1 (1 + cos phi) + K2 (1 -
// cos 2 phi) + K3 (1 + cos 3 phi); 5: sum_n C_n cos^n(phi - delta),
// n = 0..5; -1: K (phi - phi0)^2), and the gradient of phi after Blondel
// and Karplus (J. Comput. Chem. 17, 1132 (1996)):
//   d phi / d r_i = -|b2| / |m|^2 m,     d phi / d r_l = |b2| / |nn|^2 nn,
// and with s1 = b1.b2 / |b2|^2, s3 = b3.b2 / |b2|^2 (projections of the
// outer bonds on the middle one, in the directions of b1 and b3)
//   d phi / d r_j = -(1 + s1) d phi/d r_i + s3 d phi/d r_l,
//   d phi / d r_k = s1 d phi/d r_i - (1 + s3) d phi/d r_l
// (the four sum to zero); F = -dE/dphi d phi / d r, and no 1 / sin phi
// appears.  Nearly collinear
// bonds (|m|^2 |nn|^2 <= 1e-10) contribute nothing, as in the reference.
// Virial projections: sum over the atoms of (r - r_j) F.
__device__ inline void tors_body(const TorsArgs &t, int blk) {
    const ScatterArgs &a = t.s;
    const int n = blk * blockDim.x + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (n < t.n) {
        const int i = t.ti[n], j = t.tj[n], k = t.tk[n], l = t.tl[n], m = t.tpar[n];
        double b1[3] = {a.x[j] - a.x[i], a.y[j] - a.y[i], a.z[j] - a.z[i]};
        double b2[3] = {a.x[k] - a.x[j], a.y[k] - a.y[j], a.z[k] - a.z[j]};
        double b3[3] = {a.x[l] - a.x[k], a.y[l] - a.y[k], a.z[l] - a.z[k]};
        min_image(b1[0], b1[1], b1[2], a.box);
        min_image(b2[0], b2[1], b2[2], a.box);
        min_image(b3[0], b3[1], b3[2], a.box);
        auto cross = [](const double *u, const double *w, double *o) {
            o[0] = u[1] * w[2] - u[2] * w[1];
            o[1] = u[2] * w[0] - u[0] * w[2];
            o[2] = u[0] * w[1] - u[1] * w[0];
        };
        auto dot = [](const double *u, const double *w) { return u[0] * w[0] + u[1] * w[1] + u[2] * w[2]; };
        double mv[3], nv[3];
        cross(b1, b2, mv);
        cross(b2, b3, nv);
        const double mm = dot(mv, mv), nn = dot(nv, nv), bb = dot(b2, b2);
        double2 term = make_double2(0.0, 0.0);
        if (mm * nn > 1.0e-10) {
            const double lb = sqrt(bb);
            const double phi = atan2(lb * dot(b1, nv), dot(mv, nv));
            double energy = 0.0, dedp = 0.0;   // E and dE/dphi
            const int kind = t.kind[m];
            if (kind == 0) {
                const double w = t.mult[m] * phi - t.delta[m];
                energy = t.k[m] * (1.0 + cos(w));
                dedp = -t.k[m] * t.mult[m] * sin(w);
            } else if (kind == 1) {
                energy = t.k1[m] * (1.0 + cos(phi)) + t.k2[m] * (1.0 - cos(2.0 * phi)) + t.k3[m] * (1.0 + cos(3.0 * phi));
                dedp = -t.k1[m] * sin(phi) + 2.0 * t.k2[m] * sin(2.0 * phi) - 3.0 * t.k3[m] * sin(3.0 * phi);
            } else if (kind == 5) {
                const double w = phi - t.delta[m], c = cos(w);
                const double coef[6] = {t.k[m], t.k1[m], t.k2[m], t.k3[m], t.k4[m], t.k5[m]};
                double e = 0.0, de = 0.0;          // Horner in c for E and dE/dc
                for (int q = 5; q >= 0; --q) {
                    de = de * c + e;
                    e = e * c + coef[q];
                }
                energy = e;
                dedp = -sin(w) * de;
            } else if (kind == -1) {
                const double w = phi - t.delta[m];
                energy = t.k[m] * w * w;
                dedp = 2.0 * t.k[m] * w;
            }
            v[0] = energy;
            term = make_double2(fabs(phi), energy);
            const double s1 = dot(b1, b2) / bb, s3 = dot(b3, b2) / bb;
            double fi[3], fj[3], fk[3], fl[3];
            for (int q = 0; q < 3; ++q) {
                const double gi = -lb / mm * mv[q], gl = lb / nn * nv[q];
                fi[q] = -dedp * gi;
                fl[q] = -dedp * gl;
                fj[q] = -dedp * (s3 * gl - (1.0 + s1) * gi);
                fk[q] = -dedp * (s1 * gi - (1.0 + s3) * gl);
            }
            unsigned long long *F[3] = {a.fx, a.fy, a.fz};
            for (int q = 0; q < 3; ++q) {
                fx_add(F[q] + i, fi[q], a.fscale);
                fx_add(F[q] + j, fj[q], a.fscale);
                fx_add(F[q] + k, fk[q], a.fscale);
                fx_add(F[q] + l, fl[q], a.fscale);
                v[1 + q] = -b1[q] * fi[q] + b2[q] * fk[q] + (b2[q] + b3[q]) * fl[q];
            }
        }
        t.tv[n] = term;
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[4] = {Q_PINT, Q_VIRFX, Q_VIRFY, Q_VIRFZ};
    block_reduce_store<4>(v, slots, o);
}



// Per-template sums of the bonded terms (BR/BE, AR/AE, TR/TE): one block
// per template, each thread sums its strided slots in FP64, then the block
// sums in the fixed tree order of block_reduce_store.
__device__ inline void term_reduce_body(const TermArgs &a, int m) {
    double q[2] = {0.0, 0.0};
    for (int p = a.start[m] + threadIdx.x; p < a.start[m + 1]; p += blockDim.x) {
        const int s = a.slot[p];
        const double2 v = a.tv[s];
        q[0] += v.x * a.w[s];
        q[1] += v.y * a.w[s];
    }
    const int slots[2] = {2 * m, 2 * m + 1};
    block_reduce_store<2>(q, slots, a.out);
}

__global__ __launch_bounds__(256) void k_term_reduce(const TermArgs *__restrict__ args,
                                                     const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    term_reduce_body(args[m.x], m.y);
}

__device__ inline double fixd(const unsigned long long *p, int i, double inv) {
    return (double)(long long)p[i] * inv;
}

__device__ inline void combine_body(const CombineArgs &a, int blk, int group) {
    const int i = blk * blockDim.x + threadIdx.x;
    double v[2] = {0, 0};
    if (i < a.n) {
        double G[3] = {0, 0, 0}, HL[3] = {0, 0, 0}, H[3] = {0, 0, 0};
        const unsigned long long *jf = a.jf + (size_t)a.atom_slot[i] * 8;
        if (group != 1) {
            G[0] = a.flx[i] + (fixd(a.gx, i, a.inv_scale) + fixd(jf, 3, a.inv_scale));
            G[1] = a.fly[i] + (fixd(a.gy, i, a.inv_scale) + fixd(jf, 4, a.inv_scale));
            G[2] = a.flz[i] + (fixd(a.gz, i, a.inv_scale) + fixd(jf, 5, a.inv_scale));
            for (int s = 0; s < (group != 1 ? a.nslice : 0); ++s)
                for (int d = 0; d < 3; ++d) G[d] += a.recip[((size_t)s * 3 + d) * a.n + i];
        }
        if (group != 2) {
            HL[0] = a.fsx[i] + (fixd(a.hlx, i, a.inv_scale) + fixd(jf, 0, a.inv_scale));
            HL[1] = a.fsy[i] + (fixd(a.hly, i, a.inv_scale) + fixd(jf, 1, a.inv_scale));
            HL[2] = a.fsz[i] + (fixd(a.hlz, i, a.inv_scale) + fixd(jf, 2, a.inv_scale));
            H[0] = HL[0] + fixd(a.hbx, i, a.inv_scale);
            H[1] = HL[1] + fixd(a.hby, i, a.inv_scale);
            H[2] = HL[2] + fixd(a.hbz, i, a.inv_scale);
        }
        if (a.scale != 1.0) {
            // path-integral beads: physical forces / P (pimd.f:61-69)
            for (int d = 0; d < 3; ++d) { G[d] *= a.scale; HL[d] *= a.scale; H[d] *= a.scale; }
        }
        if (a.spx && group != 2) {
            H[0] += fixd(a.spx, i, a.inv_scale);
            H[1] += fixd(a.spy, i, a.inv_scale);
            H[2] += fixd(a.spz, i, a.inv_scale);
        }
        // every accumulator word of the atom back to zero, read or not
        // (a class not evaluated in this call discards what it got)
        a.gx[i] = 0ull; a.gy[i] = 0ull; a.gz[i] = 0ull;
        a.hlx[i] = 0ull; a.hly[i] = 0ull; a.hlz[i] = 0ull;
        a.hbx[i] = 0ull; a.hby[i] = 0ull; a.hbz[i] = 0ull;
        if (a.spx) { a.spx[i] = 0ull; a.spy[i] = 0ull; a.spz[i] = 0ull; }
        if (a.clear_jf)
            for (int k = 0; k < 8; ++k) a.jf[(size_t)a.atom_slot[i] * 8 + k] = 0ull;
        const int m = a.atom_mol[i];
        const double wx = a.x[i] - a.cx[m], wy = a.y[i] - a.cy[m], wz = a.z[i] - a.cz[m];
        v[0] = -G[0] * wx - G[1] * wy - G[2] * wz;
        v[1] = -HL[0] * wx - HL[1] * wy - HL[2] * wz;
        // F is the evaluated classes' sum; a class not evaluated keeps its
        // previous G or H, as FFORCES leaves GX and SLFORCES leaves HX
        // (forces.f:36-73)
        if (a.ox) {
            a.ox[i] = G[0] + H[0]; a.oy[i] = G[1] + H[1]; a.oz[i] = G[2] + H[2];
        }
        if (a.Gx && group != 1) { a.Gx[i] = G[0]; a.Gy[i] = G[1]; a.Gz[i] = G[2]; }
        if (a.Hx && group != 2) { a.Hx[i] = H[0]; a.Hy[i] = H[1]; a.Hz[i] = H[2]; }
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[2] = {Q_WIRS, Q_WIRSS};
    block_reduce_store<2>(v, slots, o);
}

// Sum rows [0, nrows) of epart, clearing each value once read (the next
// call then starts from zeros), in two launches.  k_ereduce: block blk
// takes rows kRedRows blk .. kRedRows (blk + 1) - 1; with RP = 256 / Q_N,
// thread t < RP Q_N holds slot q = t % Q_N of rows p + RP u (p = t / Q_N,
// u = 0..15: each of the block's loads covers RP consecutive rows, all
// sixteen in flight) and adds them in a pairwise tree; the block's sum per
// slot adds the RP threads' sums in order of p and goes to bpart[blk].
// k_ereduce_last (one block per system): thread (q, p) sums the block sums
// p, p + RP, ... in order, eight loads in flight, then the RP sums in order
// of p.  Every addition has a fixed order.
constexpr int kRedRP = 256 / Q_N;
static_assert(kRedRows == 16 * kRedRP, "rows per reduction block");

__device__ inline void ereduce_body(const ReduceArgs &a, int blk, double *wred) {
    const int t = threadIdx.x, q = t % Q_N, p = t / Q_N;
    const bool act = p < kRedRP;
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int r = blk * kRedRows + p + kRedRP * u;
        v[u] = act && r < a.nrows ? a.epart[(size_t)r * Q_N + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int r = blk * kRedRows + p + kRedRP * u;
        if (act && r < a.nrows) a.epart[(size_t)r * Q_N + q] = 0.0;
    }
#pragma unroll
    for (int h = 8; h > 0; h >>= 1)
#pragma unroll
        for (int u = 0; u < h; ++u) v[u] += v[u + h];
    wred[t] = v[0];
    __syncthreads();
    if (t < Q_N) {
        double b = wred[t];
        for (int k = 1; k < kRedRP; ++k) b += wred[k * Q_N + t];
        a.bpart[(size_t)blk * Q_N + t] = b;
    }
}

__device__ inline void ereduce_last_body(const ReduceArgs &a, double *wred) {
    const int t = threadIdx.x, q = t % Q_N, p = t / Q_N;
    double s = 0.0;
    if (p < kRedRP)
        for (int b0 = p; b0 < a.nblk; b0 += 8 * kRedRP) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int b = b0 + kRedRP * u;
                v[u] = b < a.nblk ? a.bpart[(size_t)b * Q_N + q] : 0.0;
            }
            s += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        }
    wred[t] = s;
    __syncthreads();
    if (t < Q_N) {
        double r = wred[t];
        for (int k = 1; k < kRedRP; ++k) r += wred[k * Q_N + t];
        a.out[t] = r;
    }
}

// Path-integral springs to both neighbouring beads (PIFORCES,
// pimd.f:321-394): E = BEADFAC m |r_k - r_k+-1|^2, force -2 BEADFAC m d.
// The neighbours' coordinates are the FP64 ones (the reference moves them
// through REAL*4 buffers, pimd.f:451-486).
__device__ inline void springs_body(const SpringArgs &a, int blk) {
    const int i = blk * blockDim.x + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};
    if (i < a.n) {
        double f[3] = {0, 0, 0};
        for (int side = 0; side < 2; ++side) {
            const double *ox = side == 0 ? a.xp : a.xm;
            const double *oy = side == 0 ? a.yp : a.ym;
            const double *oz = side == 0 ? a.zp : a.zm;
            double bx = a.x[i] - ox[i], by = a.y[i] - oy[i], bz = a.z[i] - oz[i];
            min_image(bx, by, bz, a.box);
            double bsq = bx * bx + by * by;
            bsq = bsq + bz * bz;
            const double bbb = sqrt(bsq);
            const double dbnd = bbb * a.beadfac * a.mass[i];
            const double ebdd = dbnd * bbb;
            const double forb = bbb > 0.0 ? -2.0 * dbnd / bbb : 0.0;
            f[0] += bx * forb; f[1] += by * forb; f[2] += bz * forb;
            v[0] += ebdd;
            if (side == 0) {
                // forward link only (pimd.f:336-354); VIRF* carry the
                // factor NBEADS that the caller's 1/NBEADS scaling removes
                v[1] += forb * bsq;
                v[2] += forb * bx * bx * a.nbeads;
                v[3] += forb * by * by * a.nbeads;
                v[4] += forb * bz * bz * a.nbeads;
            }
        }
        fx_add(a.fx + i, f[0], a.fscale); fx_add(a.fy + i, f[1], a.fscale); fx_add(a.fz + i, f[2], a.fscale);
        v[0] *= 0.5;     // each link is seen from both of its beads (pimd.f:387)
    }
    double *o = a.epart + (size_t)blk * Q_N;
    zero_slots(o);
    const int slots[5] = {Q_QEK, Q_VIRPI, Q_VIRFX, Q_VIRFY, Q_VIRFZ};
    block_reduce_store<5>(v, slots, o);
}

#define MDX_BATCHED(NAME, ARGS, BODY, NT)                                                    \
    __global__ __launch_bounds__(NT) void NAME(const ARGS *__restrict__ args,               \
                                               const int2 *__restrict__ map) {              \
        const int2 m = map[blockIdx.x];                                                      \
        BODY(args[m.x], m.y);                                                                \
    }

MDX_BATCHED(k_eterms, EtermArgs, eterms_body, 256)
MDX_BATCHED(k_bonds, BondArgs, bonds_body, 256)
MDX_BATCHED(k_strbnd, SpcArgs, strbnd_body, 256)
MDX_BATCHED(k_angles, AngleArgs, angles_body, 256)
MDX_BATCHED(k_tors, TorsArgs, tors_body, 256)
MDX_BATCHED(k_springs, SpringArgs, springs_body, 256)

__global__ __launch_bounds__(256) void k_combine(const CombineArgs *__restrict__ args, const int2 *__restrict__ map,
                                                 int group) {
    const int2 m = map[blockIdx.x];
    combine_body(args[m.x], m.y, group);
}

__global__ __launch_bounds__(256) void k_ereduce(const ReduceArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double wred[256];
    const int2 m = map[blockIdx.x];
    ereduce_body(args[m.x], m.y, wred);
}

__global__ __launch_bounds__(256) void k_ereduce_last(const ReduceArgs *__restrict__ args) {
    __shared__ double wred[256];
    ereduce_last_body(args[blockIdx.x], wred);
}

__global__ void k_fill_i32(int n, int *__restrict__ p, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace mdx
