// Device-resident leap-frog integrator with bond constraints and the
// Nose-Hoover thermostat of the reference's single-time-step algorithm
// (SULTAN mdstep.f:29-260, SHEJK mdstep.f:1105-1283, SCALING(3)
// mdstep.f:559-589 and 908-925, GETEMP service.f:35-92, GETCOM
// service.f:463-563).  Everything in FP64; the step-by-step order is the
// reference's:
//   OX = S; V += dt a(F); S += dt V                (k_md_kick_drift)
//   SHAKE each constrained molecule, V += dS/dt    (k_md_shake)
//   V *= 1 - SC dt (NVT); kinetic energy           (k_md_scale_ke)
//   T; SC += (T/T0 - 1) dt / QT^2                  (k_md_thermo)
//   wrap molecules by their centre of mass         (k_md_getcom)
//   [list rebuild every ICHNB steps, on the host]
//   forces at t + dt (the force-call kernels)
// Velocities are in length per time (the reference keeps momenta in
// internal units; the update V += dt F/m, S += dt V is the same map).
//
// The multiple-time-step algorithm DOUBLE (mdstep.f:281-530) uses the same
// thermostat kernels with the half long step as their step (SCALING(1):
// k_md_ke<false>, k_md_thermo, k_md_scale; SCALING(2): k_md_ke<true>,
// k_md_thermo; mdstep.f:673-679, 776-791 and 806-818, 895-901 for the
// Nose-Hoover case) and
//   V += dt/2 a(G)                                  (k_md_kick_slow)
//   nfreq times: V += dti/2 a(H); S += dti V        (k_md_kick_drift_fast)
//                GETCOM, fast (or, last, all) forces
//                V += dti/2 a(H)                    (k_md_kick_fast)
//   V += dt/2 a(G)                                  (k_md_kick_slow)
// with G the slow and H the fast forces, dti = dt / nfreq.
#include "mdx_internal.h"

namespace mdx {

__device__ inline void md_kick_drift_body(const MdArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    const double s = a.dt * a.accel * a.minv[i];
    double x = a.x[i], y = a.y[i], z = a.z[i];
    a.ox[i] = x; a.oy[i] = y; a.oz[i] = z;
    const double vx = a.vx[i] + s * a.fx[i], vy = a.vy[i] + s * a.fy[i], vz = a.vz[i] + s * a.fz[i];
    a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
    x = x + a.dt * vx; y = y + a.dt * vy; z = z + a.dt * vz;
    a.x[i] = x; a.y[i] = y; a.z[i] = z;
    a.ux[i] = x; a.uy[i] = y; a.uz[i] = z;
}

// One molecule per thread, the reference's iteration: every constraint
// whose atoms moved in the previous sweep is corrected along the old bond
// vector until |r^2 - RB^2| <= RB^2 TOLER^2 for all of them.  The molecule's
// coordinates and constraints live in registers (up to kShakeRegAtoms atoms
// and kShakeRegCons constraints, e.g. water) or are read and written in HBM
// (larger molecules).
constexpr int kShakeRegAtoms = 4;
constexpr int kShakeRegCons = 6;

// Per-constraint values that do not change between sweeps: the old bond
// vector OX(I) - OX(J) with its minimum image, 2 (1/m_I + 1/m_J), and the
// two tolerance products (the reference forms them in every sweep with the
// same operations, mdstep.f:1183-1209).
struct ShakeCons {
    double dx, dy, dz, k, tgt2, tol, bd;
};

// One constraint of a molecule held in registers, atoms I and J (slots of
// the molecule) fixed at compile time; returns whether it was corrected,
// and sets bad when the bond turned too far (VDOT < BNDSQ BDTOL).
template <int NA, int I, int J>
__device__ inline bool shake_pair(const MdArgs &a, double (&x)[NA], double (&y)[NA], double (&z)[NA],
                                  const double (&mi)[NA], const ShakeCons &c, bool &bad) {
#pragma clang fp contract(off)
    double bx = x[I] - x[J], by = y[I] - y[J], bz = z[I] - z[J];
    min_image(bx, by, bz, a.box);
    double cur2 = bx * bx + by * by;
    cur2 = cur2 + bz * bz;
    const double gap2 = c.tgt2 - cur2;
    if (!(fabs(gap2) > c.tol)) return false;
    double proj = bx * c.dx + by * c.dy;
    proj = proj + bz * c.dz;
    bad = bad || proj < c.bd;
    const double lam = gap2 / (c.k * proj);
    const double dx = c.dx * lam, dy = c.dy * lam, dz = c.dz * lam;
    const double wi = mi[I], wj = mi[J];
    x[I] = x[I] + wi * dx; y[I] = y[I] + wi * dy; z[I] = z[I] + wi * dz;
    x[J] = x[J] - wj * dx; y[J] = y[J] - wj * dy; z[J] = z[J] - wj * dz;
    return true;
}

template <int NA>
__device__ inline void md_shake_regs(const MdArgs &a, int c0, int c1, int a0) {
#pragma clang fp contract(off)
    double x[NA], y[NA], z[NA], ox[NA], oy[NA], oz[NA], mi[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int i = a0 + k;
        x[k] = a.x[i]; y[k] = a.y[i]; z[k] = a.z[i];
        ox[k] = a.ox[i]; oy[k] = a.oy[i]; oz[k] = a.oz[i];
        mi[k] = a.minv[i];
    }
    // constraints of the molecule in registers (at most kShakeRegCons), as
    // slot pairs code = 4 i + j: each constraint dispatches on its code to
    // a body with compile-time slots, so the coordinates stay in registers
    // without select chains (molecules of one type share their codes, so
    // a wavefront of them takes one branch)
    const int nc = c1 - c0;
    int code[kShakeRegCons];
    ShakeCons cs[kShakeRegCons];
    unsigned moved = 0u, moving = 0u;
#pragma unroll
    for (int c = 0; c < kShakeRegCons; ++c) {
        code[c] = 0;
        cs[c] = ShakeCons{0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0};
        if (c < nc) {
            const int2 p = a.cons[c0 + c];
            const int i = p.x - a0, j = p.y - a0;
            code[c] = 4 * i + j;
            double dx = 0.0, dy = 0.0, dz = 0.0, wi = 0.0, wj = 0.0;
#pragma unroll
            for (int k = 0; k < NA; ++k) {
                if (k == i) { dx += ox[k]; dy += oy[k]; dz += oz[k]; wi = mi[k]; }
                if (k == j) { dx -= ox[k]; dy -= oy[k]; dz -= oz[k]; wj = mi[k]; }
            }
            min_image(dx, dy, dz, a.box);
            const double t2 = a.cons_len2[c0 + c];
            cs[c] = ShakeCons{dx, dy, dz, 2.0 * (wi + wj), t2, t2 * a.tol2, t2 * a.bdtol};
            moved |= (1u << i) | (1u << j);
        }
    }
    bool done = false, bad = false;
    int iter = 0;
    while (!done && iter <= a.maxit) {
        done = true;
#pragma unroll
        for (int c = 0; c < kShakeRegCons; ++c) {
            if (c >= nc) break;
            const int i = code[c] >> 2, j = code[c] & 3;
            const unsigned bi = 1u << i, bj = 1u << j;
            if (!((moved & bi) || (moved & bj))) continue;
            bool corr = false;
            switch (code[c]) {
#define MDX_SHAKE_CASE(I, J)                                                                   \
    case 4 * I + J:                                                                            \
        if constexpr (I < NA && J < NA) corr = shake_pair<NA, I, J>(a, x, y, z, mi, cs[c], bad); \
        break;
                MDX_SHAKE_CASE(0, 1) MDX_SHAKE_CASE(1, 0) MDX_SHAKE_CASE(0, 2) MDX_SHAKE_CASE(2, 0)
                MDX_SHAKE_CASE(0, 3) MDX_SHAKE_CASE(3, 0) MDX_SHAKE_CASE(1, 2) MDX_SHAKE_CASE(2, 1)
                MDX_SHAKE_CASE(1, 3) MDX_SHAKE_CASE(3, 1) MDX_SHAKE_CASE(2, 3) MDX_SHAKE_CASE(3, 2)
#undef MDX_SHAKE_CASE
                default: break;
            }
            if (corr) {
                moving |= bi | bj;
                done = false;
            }
        }
        moved = moving;
        moving = 0u;
        ++iter;
    }
    // a bond turned too far in some sweep, or no convergence: one count
    // each for the molecule
    if (bad) atomicAdd(&a.state[MD_SHAKE_FAIL], 1.0);
    if (!done) atomicAdd(&a.state[MD_SHAKE_FAIL], 1.0);
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int i = a0 + k;
        a.x[i] = x[k]; a.y[i] = y[k]; a.z[i] = z[k];
        a.vx[i] = a.vx[i] + (x[k] - a.ux[i]) / a.dt;
        a.vy[i] = a.vy[i] + (y[k] - a.uy[i]) / a.dt;
        a.vz[i] = a.vz[i] + (z[k] - a.uz[i]) / a.dt;
    }
}

__device__ inline void md_shake_body(const MdArgs &a, int blk) {
#pragma clang fp contract(off)
    const int m = blk * 64 + threadIdx.x;
    if (m >= a.nmol) return;
    const int c0 = a.cstart[m], c1 = a.cstart[m + 1];
    if (c0 == c1) return;
    const int a0 = a.mstart[m], a1 = a.mstart[m + 1];
    if (c1 - c0 <= kShakeRegCons) {
        switch (a1 - a0) {
            case 2: md_shake_regs<2>(a, c0, c1, a0); return;
            case 3: md_shake_regs<3>(a, c0, c1, a0); return;
            case 4: md_shake_regs<4>(a, c0, c1, a0); return;
            default: break;
        }
    }
    unsigned long long moved = 0ull, moving = 0ull;   // bit (atom - a0)
    for (int c = c0; c < c1; ++c) {
        moved |= 1ull << (a.cons[c].x - a0);
        moved |= 1ull << (a.cons[c].y - a0);
    }
    bool done = false, bad = false;
    int iter = 0;
    while (!done && iter <= a.maxit) {
        done = true;
        for (int c = c0; c < c1; ++c) {
            const int i = a.cons[c].x, j = a.cons[c].y;
            const unsigned long long bi = 1ull << (i - a0), bj = 1ull << (j - a0);
            if (!((moved & bi) || (moved & bj))) continue;
            const double target2 = a.cons_len2[c];
            double bx = a.x[i] - a.x[j], by = a.y[i] - a.y[j], bz = a.z[i] - a.z[j];
            min_image(bx, by, bz, a.box);
            double cur2 = bx * bx + by * by;
            cur2 = cur2 + bz * bz;
            const double gap2 = target2 - cur2;
            if (fabs(gap2) > target2 * a.tol2) {
                double dx = a.ox[i] - a.ox[j], dy = a.oy[i] - a.oy[j], dz = a.oz[i] - a.oz[j];
                min_image(dx, dy, dz, a.box);
                double proj = bx * dx + by * dy;
                proj = proj + bz * dz;
                bad = bad || proj < target2 * a.bdtol;
                const double wi = a.minv[i], wj = a.minv[j];
                const double lam = gap2 / (2.0 * (wi + wj) * proj);
                dx = dx * lam; dy = dy * lam; dz = dz * lam;
                a.x[i] = a.x[i] + wi * dx; a.y[i] = a.y[i] + wi * dy; a.z[i] = a.z[i] + wi * dz;
                a.x[j] = a.x[j] - wj * dx; a.y[j] = a.y[j] - wj * dy; a.z[j] = a.z[j] - wj * dz;
                moving |= bi | bj;
                done = false;
            }
        }
        // atoms moved in this sweep are checked in the next one
        moved = moving;
        moving = 0ull;
        ++iter;
    }
    if (bad) atomicAdd(&a.state[MD_SHAKE_FAIL], 1.0);
    if (!done) atomicAdd(&a.state[MD_SHAKE_FAIL], 1.0);
    for (int i = a0; i < a1; ++i) {
        a.vx[i] = a.vx[i] + (a.x[i] - a.ux[i]) / a.dt;
        a.vy[i] = a.vy[i] + (a.y[i] - a.uy[i]) / a.dt;
        a.vz[i] = a.vz[i] + (a.z[i] - a.uz[i]) / a.dt;
    }
}

// Multiple time step kicks and drift (DOUBLE, mdstep.f:346-353, 364-376,
// 404-411, 423-430).
__device__ inline void md_kick_slow_body(const MdArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    const double s = 0.5 * a.dt * a.accel * a.minv[i] * a.imfac;
    a.vx[i] = a.vx[i] + s * a.gx[i];
    a.vy[i] = a.vy[i] + s * a.gy[i];
    a.vz[i] = a.vz[i] + s * a.gz[i];
}

__device__ inline void md_kick_fast_body(const MdArgs &a, int blk, bool drift) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    const double s = 0.5 * a.dti * a.accel * a.minv[i] * a.imfac;
    const double vx = a.vx[i] + s * a.hx[i], vy = a.vy[i] + s * a.hy[i], vz = a.vz[i] + s * a.hz[i];
    a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
    if (drift) {
        a.x[i] = vx * a.dti + a.x[i];
        a.y[i] = vy * a.dti + a.y[i];
        a.z[i] = vz * a.dti + a.z[i];
    }
}

// V *= 1 - SC dts (Nose-Hoover velocity scaling, mdstep.f:781-789, 812-819).
__device__ inline void md_scale_body(const MdArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n || !a.nvt) return;
    const double scv = 1.0 - a.state[MD_SC] * a.dts;
    a.vx[i] *= scv; a.vy[i] *= scv; a.vz[i] *= scv;
}

// Kinetic energy partials; with SCALE the Nose-Hoover velocity scaling
// first.
template <bool SCALE>
__device__ inline void md_scale_ke_body(const MdArgs &a, int blk, double *red) {
    const int i = blk * 256 + threadIdx.x;
    double v2 = 0.0;
    if (i < a.n) {
        double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
        if (SCALE && a.nvt) {
            const double scv = 1.0 - a.state[MD_SC] * a.dts;
            vx *= scv; vy *= scv; vz *= scv;
            a.vx[i] = vx; a.vy[i] = vy; a.vz[i] = vz;
        }
        v2 = a.mass[i] * a.mfac * (vx * vx + vy * vy + vz * vz);
    }
    red[threadIdx.x] = v2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.kepart[blk] = red[0];
}

__device__ inline void md_thermo_body(const MdArgs &a, double *red) {
    double s = 0.0;
    for (int b = threadIdx.x; b < a.nke_blocks; b += 256) s += a.kepart[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double ekin = 0.5 * a.ekin_factor * red[0];
        const double temp = 2.0 * ekin / (a.ndof * a.boltz);
        a.state[MD_EKIN] = ekin;
        a.state[MD_TEMP] = temp;
        if (a.nvt) a.state[MD_SC] = a.state[MD_SC] + (temp / a.t0 - 1.0) * a.dqt * a.dts;
    }
}

// GETCOM: S = (S - R_com) + PBC(R_com), per molecule.  A wavefront takes
// one group of consecutive molecules (at most 64, at most kComAtoms atoms
// unless one molecule is larger): their atoms pass through LDS in coalesced
// stages, each lane adds its molecule's atoms in atom order (the
// reference's serial sums), then the lanes move the atoms, lane = atom,
// with their molecule's centre and image shift from LDS.
__device__ inline void md_getcom_body(const MdArgs &a, int blk, double *sh) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    const int m0 = a.comgrp[blk], m1 = a.comgrp[blk + 1];
    const int m = m0 + lane;
    const int A0 = a.mstart[m0], A1 = a.mstart[m1];
    const int b0 = m < m1 ? a.mstart[m] : A1, b1 = m < m1 ? a.mstart[m + 1] : A1;
    double *sm_ = sh, *sx = sh + kComAtoms, *sy = sh + 2 * kComAtoms, *sz = sh + 3 * kComAtoms;
    double cx = 0.0, cy = 0.0, cz = 0.0, sm = 0.0;
    for (int c0 = A0; c0 < A1; c0 += kComAtoms) {
        const int c1 = min(c0 + kComAtoms, A1);
        for (int i = c0 + lane; i < c1; i += 64) {
            sm_[i - c0] = a.mass[i]; sx[i - c0] = a.x[i]; sy[i - c0] = a.y[i]; sz[i - c0] = a.z[i];
        }
        __syncthreads();
        for (int i = max(b0, c0); i < min(b1, c1); ++i) {
            const double mi = sm_[i - c0];
            cx = cx + mi * sx[i - c0];
            cy = cy + mi * sy[i - c0];
            cz = cz + mi * sz[i - c0];
            sm = sm + mi;
        }
        __syncthreads();
    }
    // per molecule: centre and image-shifted centre
    double *ccx = sh, *ccy = sh + 64, *ccz = sh + 128, *cpx = sh + 192, *cpy = sh + 256, *cpz = sh + 320;
    if (m < m1) {
        cx = cx / sm; cy = cy / sm; cz = cz / sm;
        double px = cx, py = cy, pz = cz;
        min_image(px, py, pz, a.box);
        ccx[lane] = cx; ccy[lane] = cy; ccz[lane] = cz;
        cpx[lane] = px; cpy[lane] = py; cpz[lane] = pz;
    }
    __syncthreads();
    // 8 atoms per lane per pass: every load of a pass before its stores
    for (int i0 = A0; i0 < A1; i0 += 8 * 64) {
        double xx[8], yy[8], zz[8];
        int kk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = min(i0 + u * 64 + lane, A1 - 1);
            xx[u] = a.x[i]; yy[u] = a.y[i]; zz[u] = a.z[i];
            kk[u] = a.atom_mol[i] - m0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * 64 + lane;
            if (i < A1) {
                const int k = kk[u];
                a.x[i] = (xx[u] - ccx[k]) + cpx[k];
                a.y[i] = (yy[u] - ccy[k]) + cpy[k];
                a.z[i] = (zz[u] - ccz[k]) + cpz[k];
            }
        }
    }
}

__device__ inline void md_trace_body(const MdArgs &a) {
    if (threadIdx.x != 0) return;
    const int step = (int)a.state[MD_STEP];
    a.state[MD_STEP] = step + 1;
    if (!a.trace || step >= a.trace_cap) return;
    const double *e = a.energy;
    double *t = a.trace + (size_t)step * kTrace;
    t[0] = e[Q_PE1] + e[Q_PE2] + e[Q_PELS1] + e[Q_PELS2] + e[Q_SPE] + a.spei + e[Q_PINT];
    t[1] = a.state[MD_EKIN];
    t[2] = a.state[MD_TEMP];
    t[3] = a.state[MD_SC];
    t[4] = e[Q_NPS] + e[Q_NPL];
}

// Drift of a short step: S += dti V (PIMD, pimd.f:122-129).
__device__ inline void md_drift_body(const MdArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    a.x[i] = a.vx[i] * a.dti + a.x[i];
    a.y[i] = a.vy[i] * a.dti + a.y[i];
    a.z[i] = a.vz[i] * a.dti + a.z[i];
}

// BEADNOSE (pimd.f:402-441) for atom i: its kinetic energy summed over the
// beads (bead mass m / nbeads; in bead order, the reference's
// MPI_ALLREDUCE), the temperature ATEMP = that / (3/2 k_B nbeads), the chain
//   xi_1 += (ATEMP / T0 - 1) dti / BTAU^2
//   xi_j += (xi_j-1^2 - FNH) dti,  xi_j-1 *= 1 - xi_j dti   (j = 2..nchain)
// and every bead's velocity of atom i scaled by 1 - xi_1 dti.
__device__ inline void pimd_beadnose_body(const PimdArgs &a, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i >= a.n) return;
    double s = 0.0;
    for (int b = 0; b < a.nbeads; ++b) {
        const double *v = a.v[b];
        const double vx = v[i], vy = v[a.n + i], vz = v[2 * (size_t)a.n + i];
        s += vx * vx + vy * vy + vz * vz;
    }
    const double atemp = a.conv * a.mass[i] * s;
    double xi_prev = a.xi[i] + (atemp / a.t0 - 1.0) * a.inv_tau2 * a.dti;
    double xi1 = xi_prev;
    for (int j = 1; j < a.nchain; ++j) {
        double xj = a.xi[(size_t)j * a.n + i];
        xj = xj + (xi_prev * xi_prev - a.fnh) * a.dti;
        xi_prev = xi_prev * (1.0 - xj * a.dti);
        a.xi[(size_t)(j - 1) * a.n + i] = xi_prev;
        if (j == 1) xi1 = xi_prev;
        xi_prev = xj;
    }
    a.xi[(size_t)(a.nchain - 1) * a.n + i] = xi_prev;
    const double scv = 1.0 - xi1 * a.dti;
    for (int b = 0; b < a.nbeads; ++b) {
        double *v = a.v[b];
        v[i] *= scv;
        v[a.n + i] *= scv;
        v[2 * (size_t)a.n + i] *= scv;
    }
}

__global__ __launch_bounds__(256) void k_pimd_beadnose(PimdArgs a) { pimd_beadnose_body(a, blockIdx.x); }

// GETCOM of path-integral beads (service.f:463-563 with PI_COM,
// pimd.f:498-520): every bead's molecular centre of mass (atoms in order),
// their average over the beads (the centroid's centre), and every bead's
// atoms of the molecule shifted by the image shift of that centroid (one
// shift per axis, PBC service.f:953-956), so a molecule's beads move
// together.  One molecule per thread.
__device__ inline void pimd_getcom_body(const PimdArgs &a, int blk) {
    const int m = blk * 256 + threadIdx.x;
    if (m >= a.nmol) return;
    const int i0 = a.mstart[m], i1 = a.mstart[m + 1];
    double sm = 0.0;
    for (int i = i0; i < i1; ++i) sm += a.mass[i];
    double c[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < a.nbeads; ++b) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int i = i0; i < i1; ++i) {
            cx = cx + a.mass[i] * a.x[b][i];
            cy = cy + a.mass[i] * a.y[b][i];
            cz = cz + a.mass[i] * a.z[b][i];
        }
        c[0] += cx / sm;
        c[1] += cy / sm;
        c[2] += cz / sm;
    }
    double sh[3];
    for (int q = 0; q < 3; ++q) {
        const double cc = c[q] / a.nbeads;
        double w = cc;
        if (fabs(w) > a.box.h[q]) w = w - copysign(a.box.l[q], w);
        sh[q] = w - cc;
    }
    // S = (S - R_b) + (R_b + shift), R_b the bead's own centre, as the
    // reference rebuilds the coordinates (service.f:523-557)
    for (int b = 0; b < a.nbeads; ++b) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int i = i0; i < i1; ++i) {
            cx = cx + a.mass[i] * a.x[b][i];
            cy = cy + a.mass[i] * a.y[b][i];
            cz = cz + a.mass[i] * a.z[b][i];
        }
        cx /= sm; cy /= sm; cz /= sm;
        const double nx = cx + sh[0], ny = cy + sh[1], nz = cz + sh[2];
        for (int i = i0; i < i1; ++i) {
            a.x[b][i] = (a.x[b][i] - cx) + nx;
            a.y[b][i] = (a.y[b][i] - cy) + ny;
            a.z[b][i] = (a.z[b][i] - cz) + nz;
        }
    }
}

__global__ __launch_bounds__(256) void k_pimd_getcom(PimdArgs a) { pimd_getcom_body(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_md_drift(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_drift_body(args[m.x], m.y);
}

__global__ __launch_bounds__(256) void k_md_kick_drift(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_kick_drift_body(args[m.x], m.y);
}

__global__ __launch_bounds__(64) void k_md_shake(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_shake_body(args[m.x], m.y);
}

template <bool SCALE>
__global__ __launch_bounds__(256) void k_md_ke(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double red[256];
    const int2 m = map[blockIdx.x];
    md_scale_ke_body<SCALE>(args[m.x], m.y, red);
}

__global__ __launch_bounds__(256) void k_md_scale(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_scale_body(args[m.x], m.y);
}

__global__ __launch_bounds__(256) void k_md_kick_slow(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_kick_slow_body(args[m.x], m.y);
}

__global__ __launch_bounds__(256) void k_md_kick_drift_fast(const MdArgs *__restrict__ args,
                                                            const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_kick_fast_body(args[m.x], m.y, true);
}

__global__ __launch_bounds__(256) void k_md_kick_fast(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_kick_fast_body(args[m.x], m.y, false);
}

__global__ __launch_bounds__(256) void k_md_thermo(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double red[256];
    const int2 m = map[blockIdx.x];
    md_thermo_body(args[m.x], red);
}

__global__ __launch_bounds__(64) void k_md_getcom(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    __shared__ double sh[4 * kComAtoms];
    const int2 m = map[blockIdx.x];
    md_getcom_body(args[m.x], m.y, sh);
}

__global__ void k_md_trace(const MdArgs *__restrict__ args, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    md_trace_body(args[m.x]);
}

}  // namespace mdx
