// k_rigid.hip -- rigid-molecule RPMD on gfx950: RATTLE (src/md_rattle.f90) and the bead-spring
// sub-steps of evolve_rig_pi (src/evolve_rigid.f90:100-278).
//
// One thread per (molecule, bead).  A molecule's three atoms and three constraints live in
// registers; the SHAKE/RATTLE iterations run in the reference's order (bonds O-H1, O-H2, H1-H2,
// repeated until no bond needs a correction), in FP64.  Molecules are independent inside a
// RATTLE call, so the per-molecule result is the reference's.  The reference stops when the
// iteration count of the LAST molecule of a call reaches itmax (src/md_rattle.f90:144-146,
// 257-259); that molecule's thread raises err[0] here, and the host turns it into RPMD_ERR_RATTLE.
#include "common.hpp"

// The RATTLE and rigid-step kernels round every product and sum once, as the reference's Fortran
// loops do when compiled without contraction: HIP's __dmul_rn and __dadd_rn are plain * and + here,
// which the compiler would otherwise fuse into FMAs.
#pragma clang fp contract(off)

namespace rpmd {

namespace {

constexpr double RAT_TOL = 1e-6;
constexpr int RAT_ITMAX = 1000;
constexpr int RAT_BI[3] = {0, 0, 1};
constexpr int RAT_BJ[3] = {1, 2, 2};

struct Mol {
    double x[3][3];  // [atom][xyz]
};

__device__ __forceinline__ void load_mol(const double* __restrict__ v, int na, int b, int m, Mol& o) {
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int j = 0; j < 3; j++) o.x[j][d] = v[((size_t)b * 3 + d) * na + 3 * m + j];
}
__device__ __forceinline__ void store_mol(double* __restrict__ v, int na, int b, int m, const Mol& o) {
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int j = 0; j < 3; j++) v[((size_t)b * 3 + d) * na + 3 * m + j] = o.x[j][d];
}

// rattle_s2 (src/md_rattle.f90:160-271) on one molecule: corrects p (and adds the constraint
// gradient to g); returns the iteration count.  vir (optional): the molecule's contributions to
// vir(1,1), (1,2), (2,2), (2,3), (3,3) and, in vlast, gdx dz of its last correction, which the
// reference's vir(1,3) = vir(1,2) - gdx dz (:232) keeps.
template <bool VIR>
__device__ __forceinline__ int rattle2_mol(const Mol& r, Mol& p, Mol& g, const double rm[3], const double rcb[3],
                                           double halfdt, double* vir, double& vlast, bool& any) {
    bool done = false;
    int it = 0;
    while (!done && it <= RAT_ITMAX) {
        done = true;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const int ii = RAT_BI[b], jj = RAT_BJ[b];
            const double dx1 = r.x[jj][0] - r.x[ii][0];
            const double dy1 = r.x[jj][1] - r.x[ii][1];
            const double dz1 = r.x[jj][2] - r.x[ii][2];
            const double rm1 = rm[jj], rm2 = rm[ii];
            const double dvx1 = p.x[jj][0] * rm1 - p.x[ii][0] * rm2;
            const double dvy1 = p.x[jj][1] * rm1 - p.x[ii][1] * rm2;
            const double dvz1 = p.x[jj][2] * rm1 - p.x[ii][2] * rm2;
            const double dot = dx1 * dvx1 + dy1 * dvy1 + dz1 * dvz1;
            const double gamma = -dot / ((rm1 + rm2) * rcb[b]);
            if (fabs(gamma) > RAT_TOL) {
                double gdx = gamma * dx1, gdy = gamma * dy1, gdz = gamma * dz1;
                p.x[jj][0] = p.x[jj][0] + gdx;
                p.x[jj][1] = p.x[jj][1] + gdy;
                p.x[jj][2] = p.x[jj][2] + gdz;
                p.x[ii][0] = p.x[ii][0] - gdx;
                p.x[ii][1] = p.x[ii][1] - gdy;
                p.x[ii][2] = p.x[ii][2] - gdz;
                gdx = gdx / halfdt;
                gdy = gdy / halfdt;
                gdz = gdz / halfdt;
                if (VIR) {
                    vir[0] = vir[0] - gdx * dx1;  // (1,1)
                    vir[1] = vir[1] - gdx * dy1;  // (1,2)
                    vlast = gdx * dz1;            // (1,3) = (1,2) - gdx dz1
                    vir[3] = vir[3] - gdy * dz1;  // (2,3)
                    vir[2] = vir[2] - gdy * dy1;  // (2,2)
                    vir[4] = vir[4] - gdz * dz1;  // (3,3)
                    any = true;
                }
                g.x[jj][0] = g.x[jj][0] + gdx;
                g.x[jj][1] = g.x[jj][1] + gdy;
                g.x[jj][2] = g.x[jj][2] + gdz;
                g.x[ii][0] = g.x[ii][0] - gdx;
                g.x[ii][1] = g.x[ii][1] - gdy;
                g.x[ii][2] = g.x[ii][2] - gdz;
                done = false;
            }
        }
        it = it + 1;
    }
    return it;
}

}  // namespace

// bead_forces (src/evolve_rigid.f90:236-278): g = m omega_n^2 (2 x_k - x_{k-1} - x_{k+1})
__global__ void k_bead_forces(const double* __restrict__ r, double* __restrict__ g, const double* __restrict__ mass,
                              int na, int nb, double omegan2) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)na * nb) return;
    const int k = (int)(t / na), a = (int)(t % na);
    const int k2 = k == 0 ? nb - 1 : k - 1, k3 = k == nb - 1 ? 0 : k + 1;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double v = 0.0;
        if (nb > 1) {
            const double dx1 = r[((size_t)k * 3 + d) * na + a], dx2 = r[((size_t)k2 * 3 + d) * na + a],
                         dx3 = r[((size_t)k3 * 3 + d) * na + a];
            v = mass[a] * omegan2 * (2.0 * dx1 - dx2 - dx3);
        }
        g[((size_t)k * 3 + d) * na + a] = v;
    }
}

// stage 1 of one bead sub-step: p -= hs g; r += p sdt / m; rattle_s1 (src/md_rattle.f90:33-158).
// The constraint gradient rattle_s1 adds to g is not kept: bead_forces overwrites g next.
__global__ void k_rig_s1(double* __restrict__ p, double* __restrict__ r, const double* __restrict__ g,
                         const double* __restrict__ mass, int na, int nb, double hs, double sdt, double rcb0,
                         double rcb1, double rcb2, int* __restrict__ err) {
    const int nm = na / 3;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nm * nb) return;
    const int b = (int)(t / nm), m = (int)(t % nm);
    Mol R, P, G, R0;
    load_mol(r, na, b, m, R);
    load_mol(p, na, b, m, P);
    load_mol(g, na, b, m, G);
    double ms[3], rm[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        ms[j] = mass[3 * m + j];
        rm[j] = 1.0 / ms[j];
    }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R0.x[j][d] = R.x[j][d];
            P.x[j][d] = P.x[j][d] - hs * G.x[j][d];
            R.x[j][d] = R.x[j][d] + P.x[j][d] * sdt / ms[j];
        }
    const double rcb[3] = {rcb0, rcb1, rcb2};
    bool done = false;
    int it = 0;
    while (!done && it < RAT_ITMAX) {
        done = true;
#pragma unroll
        for (int bd = 0; bd < 3; bd++) {
            const int ii = RAT_BI[bd], jj = RAT_BJ[bd];
            const double dx1 = R.x[jj][0] - R.x[ii][0];
            const double dy1 = R.x[jj][1] - R.x[ii][1];
            const double dz1 = R.x[jj][2] - R.x[ii][2];
            const double d2 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
            const double diff = rcb[bd] - d2;
            if (fabs(diff) > rcb[bd] * RAT_TOL) {
                const double rm1 = rm[jj], rm2 = rm[ii];
                const double dx2 = R0.x[jj][0] - R0.x[ii][0];
                const double dy2 = R0.x[jj][1] - R0.x[ii][1];
                const double dz2 = R0.x[jj][2] - R0.x[ii][2];
                const double dot = dx1 * dx2 + dy1 * dy2 + dz1 * dz2;
                const double gamma = diff / (2.0 * (rm1 + rm2) * dot);
                double gdx = gamma * dx2, gdy = gamma * dy2, gdz = gamma * dz2;
                R.x[jj][0] = R.x[jj][0] + rm1 * gdx;
                R.x[jj][1] = R.x[jj][1] + rm1 * gdy;
                R.x[jj][2] = R.x[jj][2] + rm1 * gdz;
                R.x[ii][0] = R.x[ii][0] - rm2 * gdx;
                R.x[ii][1] = R.x[ii][1] - rm2 * gdy;
                R.x[ii][2] = R.x[ii][2] - rm2 * gdz;
                gdx = gdx / sdt;
                gdy = gdy / sdt;
                gdz = gdz / sdt;
                P.x[jj][0] = P.x[jj][0] + gdx;
                P.x[jj][1] = P.x[jj][1] + gdy;
                P.x[jj][2] = P.x[jj][2] + gdz;
                P.x[ii][0] = P.x[ii][0] - gdx;
                P.x[ii][1] = P.x[ii][1] - gdy;
                P.x[ii][2] = P.x[ii][2] - gdz;
                done = false;
            }
        }
        it = it + 1;
    }
    if (m == nm - 1 && it >= RAT_ITMAX) err[0] = 1;
    store_mol(r, na, b, m, R);
    store_mol(p, na, b, m, P);
}

// stage 2 of one bead sub-step: g = bead_forces(r); p -= hs g; rattle_s2 with the sub-step
// (which also adds its constraint gradient to g, as the reference's next sub-step then uses it)
__global__ void k_rig_s2(double* __restrict__ p, const double* __restrict__ r, double* __restrict__ g,
                         const double* __restrict__ mass, int na, int nb, double hs, double sdt, double omegan2,
                         double rcb0, double rcb1, double rcb2, int* __restrict__ err) {
    const int nm = na / 3;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nm * nb) return;
    const int b = (int)(t / nm), m = (int)(t % nm);
    const int b2 = b == 0 ? nb - 1 : b - 1, b3 = b == nb - 1 ? 0 : b + 1;
    Mol R, P, G;
    load_mol(r, na, b, m, R);
    load_mol(p, na, b, m, P);
    double rm[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int a = 3 * m + j;
        const double ms = mass[a];
        rm[j] = 1.0 / ms;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double v = 0.0;
            if (nb > 1) {
                const double dx2 = r[((size_t)b2 * 3 + d) * na + a], dx3 = r[((size_t)b3 * 3 + d) * na + a];
                v = ms * omegan2 * (2.0 * R.x[j][d] - dx2 - dx3);
            }
            G.x[j][d] = v;
            P.x[j][d] = P.x[j][d] - hs * v;
        }
    }
    const double rcb[3] = {rcb0, rcb1, rcb2};
    double vl = 0.0;
    bool any = false;
    const int it = rattle2_mol<false>(R, P, G, rm, rcb, 0.5 * sdt, nullptr, vl, any);
    if (m == nm - 1 && it >= RAT_ITMAX) err[0] = 1;
    store_mol(p, na, b, m, P);
    store_mol(g, na, b, m, G);
}

// final stage of the step: p -= dt/2 dvdr; rattle_s2 with dt on every bead, adding the constraint
// gradient to dvdr (src/evolve_rigid.f90:184-200).  Per block: the sums of the five accumulated
// virial components and (highest molecule with a correction, its last gdx dz), vparts[blk][7].
__global__ __launch_bounds__(256) void k_rig_final(double* __restrict__ p, const double* __restrict__ r,
                                                   double* __restrict__ dvdr, const double* __restrict__ mass,
                                                   int na, double hdt, double dt, double rcb0, double rcb1,
                                                   double rcb2, int* __restrict__ err, double* __restrict__ vparts) {
    const int nm = na / 3;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    double vir[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double vl = 0.0;
    bool any = false;
    if (m < nm) {
        Mol R, P, G;
        load_mol(r, na, b, m, R);
        load_mol(p, na, b, m, P);
        load_mol(dvdr, na, b, m, G);
        double rm[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            rm[j] = 1.0 / mass[3 * m + j];
#pragma unroll
            for (int d = 0; d < 3; d++) P.x[j][d] = P.x[j][d] - hdt * G.x[j][d];
        }
        const double rcb[3] = {rcb0, rcb1, rcb2};
        const int it = rattle2_mol<true>(R, P, G, rm, rcb, 0.5 * dt, vir, vl, any);
        if (m == nm - 1 && it >= RAT_ITMAX) err[0] = 1;
        store_mol(p, na, b, m, P);
        store_mol(dvdr, na, b, m, G);
    }
    // five sums in fixed order
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 5; k++) vir[k] += __shfl_down(vir[k], off, 64);
    // last correcting molecule of the block: the highest index with any correction
    int who = any ? m : -1;
    for (int off = 32; off > 0; off >>= 1) {
        const int ow = __shfl_down(who, off, 64);
        const double ov = __shfl_down(vl, off, 64);
        if (ow > who) {
            who = ow;
            vl = ov;
        }
    }
    __shared__ double red[4][7];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) red[w][k] = vir[k];
        red[w][5] = (double)who;
        red[w][6] = vl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double bw = -1.0, bv = 0.0;
        const int nw = (int)(blockDim.x >> 6);
        for (int q = 0; q < nw; q++) {
            for (int k = 0; k < 5; k++) s[k] += red[q][k];
            if (red[q][5] > bw) {
                bw = red[q][5];
                bv = red[q][6];
            }
        }
        double* o = vparts + ((size_t)b * gridDim.x + blockIdx.x) * 7;
        for (int k = 0; k < 5; k++) o[k] = s[k];
        o[5] = bw;
        o[6] = bv;
    }
}

// vir_hf = (1/nb) sum_k vir_k with vir_k(1,3) = vir_k(1,2) - (last gdx dz of bead k) and the
// reference's symmetrisation (src/md_rattle.f90:264-266); out[9] column-major
__global__ void k_rig_vir(const double* __restrict__ vparts, int nblk, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++) {
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double bw = -1.0, bv = 0.0;
        for (int q = 0; q < nblk; q++) {
            const double* v = vparts + ((size_t)b * nblk + q) * 7;
            for (int k = 0; k < 5; k++) s[k] += v[k];
            if (v[5] > bw) {
                bw = v[5];
                bv = v[6];
            }
        }
        const double v11 = s[0], v12 = s[1], v22 = s[2], v23 = s[3], v33 = s[4];
        const double v13 = bw >= 0.0 ? v12 - bv : 0.0;
        const double w[9] = {v11, v12, v13, v12, v22, v23, v13, v23, v33};  // column-major, symmetric
        for (int k = 0; k < 9; k++) acc[k] = acc[k] + w[k];
    }
    for (int k = 0; k < 9; k++) out[k] = acc[k] / (double)nb;
}

int rigid_blocks(int na) { return (na / 3 + 255) / 256; }

void launch_bead_forces(const double* r, double* g, const double* mass, int na, int nb, double omegan2,
                        hipStream_t s) {
    const size_t n = (size_t)na * nb;
    k_bead_forces<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(r, g, mass, na, nb, omegan2);
}
void launch_rig_substep(double* p, double* r, double* g, const double* mass, int na, int nb, double hs, double sdt,
                        double omegan2, const double rcb[3], int* err, hipStream_t s) {
    const size_t n = (size_t)(na / 3) * nb;
    const unsigned grid = (unsigned)((n + 255) / 256);
    k_rig_s1<<<grid, 256, 0, s>>>(p, r, g, mass, na, nb, hs, sdt, rcb[0], rcb[1], rcb[2], err);
    k_rig_s2<<<grid, 256, 0, s>>>(p, r, g, mass, na, nb, hs, sdt, omegan2, rcb[0], rcb[1], rcb[2], err);
}
void launch_rig_final(double* p, const double* r, double* dvdr, const double* mass, int na, int nb, double hdt,
                      double dt, const double rcb[3], int* err, double* vparts, double* vir_out, hipStream_t s) {
    const dim3 grid(rigid_blocks(na), nb);
    k_rig_final<<<grid, 256, 0, s>>>(p, r, dvdr, mass, na, hdt, dt, rcb[0], rcb[1], rcb[2], err, vparts);
    k_rig_vir<<<1, 64, 0, s>>>(vparts, (int)grid.x, nb, vir_out);
}

}  // namespace rpmd
