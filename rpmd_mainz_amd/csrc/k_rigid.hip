// k_rigid.hip -- rigid-molecule RPMD on gfx950: RATTLE (src/md_rattle.f90) and the bead-spring
// sub-steps of evolve_rig_pi (src/evolve_rigid.f90:100-278).
//
// One thread per (molecule, bead).  A molecule's three atoms and three constraints live in
// registers; the sweeps over the constraints run in the reference's order (O-H1, O-H2, H1-H2,
// repeated until no constraint needs a correction), in FP64, with the tolerance a parameter.  Molecules are independent inside a
// RATTLE call, so the per-molecule result is the reference's.  The reference stops when the
// iteration count of the LAST molecule of a call reaches itmax (src/md_rattle.f90:144-146,
// 257-259); that molecule's thread raises err[0] here, and the host turns it into RPMD_ERR_RATTLE.
#include <algorithm>

#include "common.hpp"

// The RATTLE and rigid-step kernels round every product and sum once, as the reference's Fortran
// loops do when compiled without contraction: HIP's __dmul_rn and __dadd_rn are plain * and + here,
// which the compiler would otherwise fuse into FMAs.
#pragma clang fp contract(off)

namespace rpmd {

namespace {

constexpr int RAT_ITMAX = 1000;
// the three constraints of a molecule (O, H1, H2): O-H1, O-H2, H1-H2, in the reference's order
constexpr int CON_A[3] = {0, 0, 1};
constexpr int CON_B[3] = {1, 2, 2};

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
// (x + y) + z of the componentwise products, the reference's summation order
__device__ __forceinline__ double dot3(const double u[3], const double w[3]) {
    return u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
}

// Velocity stage of RATTLE (rattle_s2, src/md_rattle.f90:160-271) on one molecule: the relative
// velocity of every constrained pair along its bond is removed by the multiplier
// lam = -(u . (p_b/m_b - p_a/m_a)) / ((1/m_a + 1/m_b) |u_0|^2), |u_0| the bond length, until every
// |lam| <= tol; p is corrected and the constraint force lam u / (dt/2) added to g.  Returns the
// sweep count.  VIR: the molecule's contributions to vir(1,1), (1,2), (2,2), (2,3), (3,3) and, in
// vlast, the x-z term of its last correction, which the reference's vir(1,3) = vir(1,2) - ... keeps
// (:232).
template <bool VIR>
__device__ __forceinline__ int rattle_vel(const Mol& r, Mol& p, Mol& g, const double inv_m[3], const double len2[3],
                                          double halfdt, double tol, double* vir, double& vlast, bool& any) {
    bool settled = false;
    int sweeps = 0;
    while (!settled && sweeps <= RAT_ITMAX) {
        settled = true;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int ia = CON_A[c], ib = CON_B[c];
            double u[3], dv[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                u[d] = r.x[ib][d] - r.x[ia][d];
                dv[d] = p.x[ib][d] * inv_m[ib] - p.x[ia][d] * inv_m[ia];
            }
            const double lam = -dot3(u, dv) / ((inv_m[ib] + inv_m[ia]) * len2[c]);
            if (fabs(lam) > tol) {
                double f[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    f[d] = lam * u[d];
                    p.x[ib][d] = p.x[ib][d] + f[d];
                    p.x[ia][d] = p.x[ia][d] - f[d];
                }
#pragma unroll
                for (int d = 0; d < 3; d++) f[d] = f[d] / halfdt;
                if (VIR) {
                    vir[0] = vir[0] - f[0] * u[0];  // (1,1)
                    vir[1] = vir[1] - f[0] * u[1];  // (1,2)
                    vlast = f[0] * u[2];            // (1,3) = (1,2) - this
                    vir[3] = vir[3] - f[1] * u[2];  // (2,3)
                    vir[2] = vir[2] - f[1] * u[1];  // (2,2)
                    vir[4] = vir[4] - f[2] * u[2];  // (3,3)
                    any = true;
                }
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    g.x[ib][d] = g.x[ib][d] + f[d];
                    g.x[ia][d] = g.x[ia][d] - f[d];
                }
                settled = false;
            }
        }
        sweeps = sweeps + 1;
    }
    return sweeps;
}

}  // namespace

// All nsub bead sub-steps of a rigid step in one launch (src/evolve_rigid.f90:131-170): a block
// owns every bead of MB molecules, thread (molecule m = t % MB, bead b = t / MB).  A molecule's
// positions, momenta and gradient stay in registers through the sub-steps; the positions go through
// LDS, where the bead forces read the neighbouring beads after a barrier.  The gradient at entry is
// bead_forces (src/evolve_rigid.f90:236-278: g = m omega_n^2 (2 x_k - x_{k-1} - x_{k+1})) of the
// entry positions.  Per sub-step:
//   stage 1: p -= hs g; r += p sdt / m; then the position stage of RATTLE (rattle_s1,
//     src/md_rattle.f90:33-158): per constrained pair with squared length off its target by more
//     than tol (relative), the pair moves along its bond vector u_0 at the start of the sub-step by
//     lam = (|u_0|^2_target - |u|^2) / (2 (1/m_a + 1/m_b) u . u_0), weighted by the inverse masses,
//     and the momenta take lam u_0 / sdt (the constraint gradient rattle_s1 adds to g is not kept:
//     bead_forces overwrites g next);
//   barrier;
//   stage 2: g = bead_forces(r); p -= hs g; rattle_s2 with the sub-step, which adds its constraint
//     gradient to g (the next sub-step's stage 1 uses it, as the reference's does);
//   barrier.
__global__ __launch_bounds__(256) void k_rig_substeps(double* __restrict__ p, double* __restrict__ r,
                                                      const double* __restrict__ mass, int na, int nb, int mb,
                                                      int nsub, double hs, double sdt, double omegan2, double len2_0,
                                                      double len2_1, double len2_2, double tol,
                                                      int* __restrict__ err) {
    extern __shared__ double lds_r[];  // [bead][molecule of the block][atom][xyz]
    const int nm = na / 3;
    const int mloc = threadIdx.x % mb, b = threadIdx.x / mb;
    const int m = blockIdx.x * mb + mloc;
    const bool live = m < nm && b < nb;
    const int b2 = b == 0 ? nb - 1 : b - 1, b3 = b == nb - 1 ? 0 : b + 1;
    double* mine = lds_r + ((size_t)b * mb + mloc) * 9;
    const double* left = lds_r + ((size_t)b2 * mb + mloc) * 9;
    const double* right = lds_r + ((size_t)b3 * mb + mloc) * 9;
    const double len2[3] = {len2_0, len2_1, len2_2};
    Mol R, P, G;
    double ms[3], inv_m[3];
    if (live) {
        load_mol(r, na, b, m, R);
        load_mol(p, na, b, m, P);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            ms[j] = mass[3 * m + j];
            inv_m[j] = 1.0 / ms[j];
        }
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) mine[3 * j + d] = R.x[j][d];
    }
    __syncthreads();
    // bead_forces of the entry positions
    if (live) {
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double v = 0.0;
                if (nb > 1) {
                    const double dx1 = R.x[j][d], dx2 = left[3 * j + d], dx3 = right[3 * j + d];
                    v = ms[j] * omegan2 * (2.0 * dx1 - dx2 - dx3);
                }
                G.x[j][d] = v;
            }
    }
    for (int is = 0; is < nsub; is++) {
        __syncthreads();  // every read of the positions in LDS is done
        if (live) {
            // stage 1
            Mol R0;
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    R0.x[j][d] = R.x[j][d];
                    P.x[j][d] = P.x[j][d] - hs * G.x[j][d];
                    R.x[j][d] = R.x[j][d] + P.x[j][d] * sdt / ms[j];
                }
            bool settled = false;
            int sweeps = 0;
            while (!settled && sweeps < RAT_ITMAX) {
                settled = true;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int ia = CON_A[c], ib = CON_B[c];
                    double u[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) u[d] = R.x[ib][d] - R.x[ia][d];
                    const double miss = len2[c] - dot3(u, u);
                    if (fabs(miss) > len2[c] * tol) {
                        double u0[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) u0[d] = R0.x[ib][d] - R0.x[ia][d];
                        const double lam = miss / (2.0 * (inv_m[ib] + inv_m[ia]) * dot3(u, u0));
                        double f[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) f[d] = lam * u0[d];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            R.x[ib][d] = R.x[ib][d] + inv_m[ib] * f[d];
                            R.x[ia][d] = R.x[ia][d] - inv_m[ia] * f[d];
                        }
#pragma unroll
                        for (int d = 0; d < 3; d++) f[d] = f[d] / sdt;
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            P.x[ib][d] = P.x[ib][d] + f[d];
                            P.x[ia][d] = P.x[ia][d] - f[d];
                        }
                        settled = false;
                    }
                }
                sweeps = sweeps + 1;
            }
            if (m == nm - 1 && sweeps >= RAT_ITMAX) err[0] = 1;
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int d = 0; d < 3; d++) mine[3 * j + d] = R.x[j][d];
        }
        __syncthreads();  // every bead's new positions are in LDS
        if (live) {
            // stage 2
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double v = 0.0;
                    if (nb > 1) {
                        const double dx2 = left[3 * j + d], dx3 = right[3 * j + d];
                        v = ms[j] * omegan2 * (2.0 * R.x[j][d] - dx2 - dx3);
                    }
                    G.x[j][d] = v;
                    P.x[j][d] = P.x[j][d] - hs * v;
                }
            double vl = 0.0;
            bool any = false;
            const int sweeps = rattle_vel<false>(R, P, G, inv_m, len2, 0.5 * sdt, tol, nullptr, vl, any);
            if (m == nm - 1 && sweeps >= RAT_ITMAX) err[0] = 1;
        }
    }
    if (live) {
        store_mol(r, na, b, m, R);
        store_mol(p, na, b, m, P);
    }
}

// final stage of the step: p -= dt/2 dvdr; rattle_s2 with dt on every bead, adding the constraint
// gradient to dvdr (src/evolve_rigid.f90:184-200).  Per block: the sums of the five accumulated
// virial components and (highest molecule with a correction, its last gdx dz), vparts[blk][7].
__global__ __launch_bounds__(256) void k_rig_final(double* __restrict__ p, const double* __restrict__ r,
                                                   double* __restrict__ dvdr, const double* __restrict__ mass,
                                                   int na, double hdt, double dt, double len2_0, double len2_1,
                                                   double len2_2, double tol, int* __restrict__ err,
                                                   double* __restrict__ vparts) {
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
        double inv_m[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            inv_m[j] = 1.0 / mass[3 * m + j];
#pragma unroll
            for (int d = 0; d < 3; d++) P.x[j][d] = P.x[j][d] - hdt * G.x[j][d];
        }
        const double len2[3] = {len2_0, len2_1, len2_2};
        const int sweeps = rattle_vel<true>(R, P, G, inv_m, len2, 0.5 * dt, tol, vir, vl, any);
        if (m == nm - 1 && sweeps >= RAT_ITMAX) err[0] = 1;
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
__global__ __launch_bounds__(256) void k_rig_vir(const double* __restrict__ vparts, int nblk, int nb,
                                                 double* __restrict__ out) {
    // thread b: bead b's virial from its blocks' parts in block order; thread 0 then sums the beads in
    // bead order (the order of one serial loop over beads and blocks)
    __shared__ double wb[256][9];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
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
        for (int k = 0; k < 9; k++) wb[b][k] = w[k];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < 9; k++) acc[k] = acc[k] + wb[b][k];
    for (int k = 0; k < 9; k++) out[k] = acc[k] / (double)nb;
}

// The velocity stage of RATTLE alone on every (molecule, bead): after a thermostat kick, the
// momenta are put back on the constraint surface's tangent space (no gradient, no virial).
// scaled: p holds mass-scaled momenta p / sqrt(m) (a GLE auxiliary channel), projected the same way.
__global__ void k_rig_project(double* __restrict__ p, const double* __restrict__ r, const double* __restrict__ mass,
                              int na, int nb, double len2_0, double len2_1, double len2_2, double tol, int scaled,
                              int* __restrict__ err) {
    const int nm = na / 3;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nm * nb) return;
    const int b = (int)(t / nm), m = (int)(t % nm);
    Mol R, P, G{};
    load_mol(r, na, b, m, R);
    load_mol(p, na, b, m, P);
    double inv_m[3], sq[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        inv_m[j] = 1.0 / mass[3 * m + j];
        sq[j] = sqrt(mass[3 * m + j]);
    }
    if (scaled)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) P.x[j][d] = P.x[j][d] * sq[j];
    const double len2[3] = {len2_0, len2_1, len2_2};
    double vl = 0.0;
    bool any = false;
    const int sweeps = rattle_vel<false>(R, P, G, inv_m, len2, 1.0, tol, nullptr, vl, any);
    if (m == nm - 1 && sweeps >= RAT_ITMAX) err[0] = 1;
    if (scaled)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) P.x[j][d] = P.x[j][d] / sq[j];
    store_mol(p, na, b, m, P);
}

int rigid_blocks(int na) { return (na / 3 + 255) / 256; }

// the molecules of a block: every bead of each (nb <= 256), about one block per CU when the molecules
// allow, at most 256 threads
int rig_substep_mols(int nm, int nb) {
    const int cap = std::max(1, 256 / nb);
    return std::max(1, std::min(cap, nm / 256));
}
void launch_rig_substeps(double* p, double* r, const double* mass, int na, int nb, int nsub, double hs, double sdt,
                         double omegan2, const double len2[3], double tol, int* err, hipStream_t s) {
    const int nm = na / 3, mb = rig_substep_mols(nm, nb);
    const unsigned grid = (unsigned)((nm + mb - 1) / mb);
    const size_t lds = (size_t)mb * nb * 9 * sizeof(double);
    k_rig_substeps<<<grid, mb * nb, lds, s>>>(p, r, mass, na, nb, mb, nsub, hs, sdt, omegan2, len2[0], len2[1],
                                              len2[2], tol, err);
}
void launch_rig_project(double* p, const double* r, const double* mass, int na, int nb, const double len2[3],
                        double tol, int scaled, int* err, hipStream_t s) {
    const size_t n = (size_t)(na / 3) * nb;
    k_rig_project<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, r, mass, na, nb, len2[0], len2[1], len2[2], tol,
                                                            scaled, err);
}
void launch_rig_final(double* p, const double* r, double* dvdr, const double* mass, int na, int nb, double hdt,
                      double dt, const double len2[3], double tol, int* err, double* vparts, double* vir_out,
                      hipStream_t s) {
    const dim3 grid(rigid_blocks(na), nb);
    k_rig_final<<<grid, 256, 0, s>>>(p, r, dvdr, mass, na, hdt, dt, len2[0], len2[1], len2[2], tol, err, vparts);
    k_rig_vir<<<1, 256, 0, s>>>(vparts, (int)grid.x, nb, vir_out);
}

}  // namespace rpmd
