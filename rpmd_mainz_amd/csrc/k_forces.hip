// k_forces.hip -- force kernels of the RPMD path on gfx950.
//
// All bead-batched: one launch covers every bead of a forces() call (grid.y or grid.z = bead).
// Per-atom gradients are owned by one thread per atom (full neighbour sums, no atomics), so every
// result is deterministic run to run.  Energies and virials are accumulated in FP64 per thread,
// reduced per block into parts[block][10] and summed in a fixed order by k_reduce_parts.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace rpmd {

// A&S 7.1.26 erfc constants (src/ewald_cubic.f90:94-99)
constexpr double AS_P = 3.0525860, AS_A1 = 0.254829592, AS_A2 = -0.284496736, AS_A3 = 1.421413741,
                 AS_A4 = -1.453152027, AS_A5 = 1.061405429;

// ------------------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* out) {
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < NV; i++) v[i] += __shfl_down(v[i], off, 64);
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int k = 0; k < nw; k++) s += red[k][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// out[k] = scale * sum_p parts[p*10 + k], k < 10, for up to RED_JOBS part ranges in one launch; one
// block of 256 threads per (k, job), fixed order
__device__ __forceinline__ void reduce_parts_block(const ReduceJobs& jobs, double* out, int k, int jb) {
    __shared__ double red[256];
    const double* parts = jobs.parts[jb];
    const int nparts = jobs.nparts[jb];
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += 256) s += parts[(size_t)p * 10 + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[jobs.out[jb] + k] = jobs.scale[jb] * red[0];
}
__global__ __launch_bounds__(256) void k_reduce_parts(ReduceJobs jobs, double* out) {
    reduce_parts_block(jobs, out, blockIdx.x, blockIdx.y);
}

void launch_reduce_ev(const ReduceJobs& jobs, double* out, hipStream_t s) {
    if (jobs.n > 0) k_reduce_parts<<<dim3(10, jobs.n), 256, 0, s>>>(jobs, out);
}

// ------------------------------------------------------------------------------------------------
// layout: Fortran r(3,na,nb) <-> SoA X[b][d][a]
// ------------------------------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, int na, int nb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)na * nb;
    if (i >= n) return;
    size_t b = i / na, a = i % na;
#pragma unroll
    for (int d = 0; d < 3; d++) soa[(b * 3 + d) * na + a] = aos[i * 3 + d];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, int na, int nb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)na * nb;
    if (i >= n) return;
    size_t b = i / na, a = i % na;
#pragma unroll
    for (int d = 0; d < 3; d++) aos[i * 3 + d] = soa[(b * 3 + d) * na + a];
}
void launch_aos_to_soa(const double* aos, double* soa, int na, int nb, hipStream_t s) {
    size_t n = (size_t)na * nb;
    k_aos_to_soa<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(aos, soa, na, nb);
}
void launch_soa_to_aos(const double* soa, double* aos, int na, int nb, hipStream_t s) {
    size_t n = (size_t)na * nb;
    k_soa_to_aos<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(soa, aos, na, nb);
}

// M-site substitution r_M = alpha r_O + (1-alpha)/2 (r_H1 + r_H2) (src/md_forces.f90:181-188),
// optionally scaled to the unit box (src/ewald_cubic.f90:21-22).
__global__ void k_sites(const double* __restrict__ pos, double* __restrict__ site, int na, int nb, double alpha,
                        double s0, double s1, double s2) {
    int nm = na / 3;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nm * nb) return;
    size_t b = i / nm, m = i % nm;
    double alpha2 = 0.5 * (1.0 - alpha);
    double sc[3] = {s0, s1, s2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double* p = pos + (b * 3 + d) * na + 3 * m;
        double* q = site + (b * 3 + d) * na + 3 * m;
        double o = alpha * p[0] + alpha2 * (p[1] + p[2]);
        q[0] = sc[d] * o;
        q[1] = sc[d] * p[1];
        q[2] = sc[d] * p[2];
    }
}
void launch_sites(const double* pos, double* site, int na, int nb, double alpha, int unit, Box box, hipStream_t s) {
    size_t n = (size_t)(na / 3) * nb;
    double s0 = unit ? box.inv[0] : 1.0, s1 = unit ? box.inv[1] : 1.0, s2 = unit ? box.inv[2] : 1.0;
    k_sites<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos, site, na, nb, alpha, s0, s1, s2);
}

// grad (+)= scale * chain rule of the M-site (src/md_forces.f90:206-215) applied to the site gradient,
// the sum of npart partial site gradients (parts[p*pstride + ...]) taken in fixed order
// lj (optional): the oxygens' O-O gradient per molecule [bead][dim][molecule], added last.
// kp (optional): then the momentum kick of the three atoms, p = (p - a1 g1) - a2 grad with each
// product and difference rounded as k_kick does, and the noise counter's step advance
// nchain: the blocks of the chain rule; blocks past them run the energy reductions red (component
// (b - nchain) % 10, job (b - nchain) / 10) into ev_out
__global__ __launch_bounds__(256) void k_chain(double* __restrict__ grad, const double* __restrict__ parts, int npart,
                                               size_t pstride, int na, int nb, double alpha, double scale, int acc,
                                               const double* __restrict__ lj, double* __restrict__ kp,
                                               const double* __restrict__ kg1, double ka1, double ka2,
                                               uint64_t* __restrict__ ctr, int nchain, ReduceJobs red,
                                               double* __restrict__ ev_out) {
    if ((int)blockIdx.x >= nchain) {
        const int r = blockIdx.x - nchain;
        reduce_parts_block(red, ev_out, r % 10, r / 10);
        return;
    }
    int nm = na / 3;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ctr && i == 0) ctrl_add(ctr, (uint64_t)RPMD_STREAMS_PER_STEP);
    if (i >= (size_t)nm * nb) return;
    size_t b = i / nm, m = i % nm;
    double alpha2 = 0.5 * (1.0 - alpha);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        size_t o = (b * 3 + d) * na + 3 * m;
        double s0 = parts[o], s1 = parts[o + 1], s2 = parts[o + 2];
        for (int p = 1; p < npart; p++) {
            s0 += parts[p * pstride + o];
            s1 += parts[p * pstride + o + 1];
            s2 += parts[p * pstride + o + 2];
        }
        double g0 = scale * s0, g1 = scale * s1, g2 = scale * s2;
        const double b0 = acc ? grad[o] : 0.0, b1 = acc ? grad[o + 1] : 0.0, b2 = acc ? grad[o + 2] : 0.0;
        const double r0 = lj ? (b0 + alpha * g0) + lj[(b * 3 + d) * nm + m] : b0 + alpha * g0;
        const double r1 = b1 + (g1 + alpha2 * g0), r2 = b2 + (g2 + alpha2 * g0);
        grad[o] = r0;
        grad[o + 1] = r1;
        grad[o + 2] = r2;
        if (kp) {
            const double rr[3] = {r0, r1, r2};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double v = kp[o + k];
                v = rn_add(v, -rn_mul(ka1, kg1[o + k]));
                v = rn_add(v, -rn_mul(ka2, rr[k]));
                kp[o + k] = v;
            }
        }
    }
}
// acc = 0: the first writer of grad in a forces call (stores instead of adding)
void launch_chain_rule(double* grad, const double* parts, int npart, int na, int nb, double alpha, double scale,
                       int acc, hipStream_t s, const double* lj, const KickArgs* kick) {
    size_t n = (size_t)(na / 3) * nb;
    const int nchain = (int)((n + 255) / 256);
    ReduceJobs red;
    if (kick && kick->red) red = *kick->red;
    k_chain<<<(unsigned)(nchain + 10 * red.n), 256, 0, s>>>(
        grad, parts, npart, (size_t)3 * na * nb, na, nb, alpha, scale, acc, lj, kick ? kick->p : nullptr,
        kick ? kick->g1 : nullptr, kick ? kick->a1 : 0.0, kick ? kick->a2 : 0.0, kick ? kick->ctr : nullptr, nchain,
        red, kick ? kick->ev_out : nullptr);
}

// bead average (src/md_forces.f90:39-44): sum over beads in order, then divide
__global__ void k_centroid(const double* __restrict__ pos, double* __restrict__ cen, int na, int nb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)3 * na) return;
    size_t d = i / na, a = i % na;
    double s = 0.0;
#pragma unroll 8
    for (int b = 0; b < nb; b++) s = s + pos[((size_t)b * 3 + d) * na + a];
    cen[d * na + a] = s / (double)nb;
}
void launch_centroid(const double* pos, double* cen, int na, int nb, hipStream_t s) {
    size_t n = (size_t)3 * na;
    k_centroid<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos, cen, na, nb);
}
// centroids as k_centroid (same sums), and for the oxygens (atoms 0, 3, 6, ...) the largest squared
// displacement of one centroid coordinate from cb[] (its value at the O-O list build), max-reduced
// into dmax as bit patterns (NaN reference positions count as huge); |dc|^2 <= 3 max_d dc_d^2
__global__ __launch_bounds__(256) void k_centroid_vl(const double* __restrict__ pos, double* __restrict__ cen, int na,
                                                     int nb, const double* __restrict__ cb,
                                                     unsigned long long* __restrict__ dmax) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < (size_t)3 * na) {
        const size_t d = i / na, a = i % na;
        double s = 0.0;
#pragma unroll 8
        for (int b = 0; b < nb; b++) s = s + pos[((size_t)b * 3 + d) * na + a];
        const double c = s / (double)nb;
        cen[d * na + a] = c;
        if (a % 3 == 0) {
            const double e = c - cb[d * (na / 3) + a / 3];
            m = e * e;
            if (!(m <= 1e300)) m = 1e300;
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        ctrl_max(dmax, (unsigned long long)__double_as_longlong(fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))));
}
void launch_centroid_vl(const double* pos, double* cen, int na, int nb, const double* cb, unsigned long long* dmax,
                        hipStream_t s) {
    const size_t n = (size_t)3 * na;
    k_centroid_vl<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos, cen, na, nb, cb, dmax);
}

__global__ void k_axpy(double* y, const double* x, double a, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
void launch_axpy(double* y, const double* x, double a, size_t n, hipStream_t s) {
    k_axpy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, x, a, n);
}
__global__ void k_zero(double* y, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}
void launch_zero(double* y, size_t n, hipStream_t s) {
    if (n) k_zero<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, n);
}
__global__ void k_zero_int(int* y, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0;
}
void launch_zero_int(int* y, size_t n, hipStream_t s) {
    if (n) k_zero_int<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, n);
}
__global__ void k_fill32(uint32_t* y, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}
void launch_fill32(void* y, uint32_t v, size_t nwords, hipStream_t s) {
    if (nwords) k_fill32<<<(unsigned)((nwords + 255) / 256), 256, 0, s>>>((uint32_t*)y, v, nwords);
}

// ------------------------------------------------------------------------------------------------
// cell coordinates of the Ewald sites; the cells are numbered in Morton (Z-curve) order over a
// power-of-two grid, so that sites sorted by cell stay close in space
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t part1by2(uint32_t x) {
    x &= 0x3ff;
    x = (x | (x << 16)) & 0x030000FF;
    x = (x | (x << 8)) & 0x0300F00F;
    x = (x | (x << 4)) & 0x030C30C3;
    x = (x | (x << 2)) & 0x09249249;
    return x;
}

__device__ __forceinline__ int cell_coord(double x, double inv, int nc) {
    double f = x * inv;
    f = f - floor(f);
    int c = (int)(f * nc);
    return c < nc ? (c >= 0 ? c : 0) : nc - 1;
}

// ------------------------------------------------------------------------------------------------
// real-space Ewald with the A&S erfc (src/ewald_cubic.f90:78-233; src/ewald_noncubic.f90:54-312)
//
// Per-site Verlet lists with a skin, one list for all beads.  At the build, site k has the bead
// centroid c_k and the spread s_k (the largest distance of one of its beads from c_k); the list of k
// holds the sites j with |c_k - c_j| < rc + skin + s_k + s_j.  It stays exact -- it contains every
// pair of any bead now within rc -- while every bead of every site stays within s_k + skin/2 of its
// build centroid (then |x_kb - x_jb| >= |c_k - c_j| - s_k - s_j - skin), however the ring polymers
// move inside those spheres.  Each call measures the largest excursion beyond s_k on the device
// (k_sites_vl, an atomic max); the build kernels return at once unless it exceeds skin/2 (VGate), so
// a captured hipGraph rebuilds exactly when needed.  The build (k_vl_sort) sorts the sites by cell in Morton
// order, groups runs of 64 sorted sites into clusters with bounding boxes and lists the cluster pairs
// within reach; k_vl_build then fills each site's list (sorted indices, cluster order).  The pair
// kernel gives VLS lanes to a sorted site, which walk its list: every pair is evaluated from both
// ends (energy and virial with a factor 1/2) and each site's gradient is summed in a fixed order, so
// results are deterministic.  T is the pair-body type; in FP32 the distances come from 32-bit
// fixed-point fractional coordinates, the pair functions from a table (PairTab), and the sums are
// flushed into FP64 every 64 pairs of a lane.
// ------------------------------------------------------------------------------------------------
// FP64: du = erfc_AS(alpha r)/r (erfc - 1 within a molecule), dur = (sfac e^{-x^2} - du)/r^2, the
// reference's expressions (src/ewald_cubic.f90:126-140) with the two divisions by r and r^2 taken as
// products with 1/r (rounding-level differences only)
template <typename T>
__device__ __forceinline__ void erfc_pair(T dr2, T alpha, T sfac, bool same, T& du, T& dur) {
    const T rinv = (T)1 / sqrt(dr2);
    const T dr = dr2 * rinv;
    T x = alpha * dr;
    T e = exp(-x * x);
    T t = (T)AS_P / ((T)AS_P + x);
    T erfc_ = e * t * ((T)AS_A1 + t * ((T)AS_A2 + t * ((T)AS_A3 + t * ((T)AS_A4 + t * (T)AS_A5))));
    if (same) erfc_ = erfc_ - (T)1;
    du = erfc_ * rinv;
    dur = (sfac * e - du) * (rinv * rinv);
}

#define CLU 64

__device__ __forceinline__ double wrap01(double f) { return f - floor(f); }
// a wrapped fractional coordinate in 32-bit fixed point, low bit clear: differences of two in 32-bit
// integer arithmetic are the minimum image (mod 2^32) exactly
__device__ __forceinline__ uint32_t frac_fixed(double f) {
    return (uint32_t)fmin(wrap01(f) * 4294967296.0, 4294967295.0) & ~1u;
}
__device__ __forceinline__ uint4 site_fixed(double x, double y, double z, double q, double f0, double f1, double f2) {
    return make_uint4(frac_fixed(x * f0), frac_fixed(y * f1), frac_fixed(z * f2), __float_as_uint((float)q));
}

// Per site of every bead: the Ewald site (the M-site combination of the oxygen and its hydrogens at
// a%3 == 0, as k_sites), scaled by sc; its phase factors e^{i lat_d x_d}; how far it has left the
// sphere of the list build around the site's bead centroid c_k with the site's bead spread s_k as
// radius, e = max(0, |x_b - c_k| - s_k), squared and max-reduced into dmax as bit patterns
// (non-negative doubles order as their bits; the NaN centroids of a list never built count as huge):
// while every bead stays within s_k + skin / 2 of c_k, every pair within rc is on the list (it lists
// the pairs with |c_k - c_j| < rc + skin + s_k + s_j), however the ring polymers move inside their
// spheres.  When gather is set, the site written at its sorted index iperm[] (the same for every
// bead) as (x, y, z, q) in FP64 and as wrapped fractional FP32 coordinates (inverse box side f).
// Thread per (site, bead), a wave over 64 consecutive sites of one bead (its loads and stores of
// positions, sites and phases contiguous), a block over four beads; the centroids and spreads that a
// build needs are formed by k_vl_cu, only when the build runs.
__global__ __launch_bounds__(256) void k_sites_vl(const double* __restrict__ pos, double* __restrict__ site, int na,
                                                  int nb, double alpha, double s0, double s1, double s2, double l0,
                                                  double l1, double l2, double2* __restrict__ ph,
                                                  const double* __restrict__ q, const double* __restrict__ cb,
                                                  const int* __restrict__ iperm, int gather, double f0, double f1,
                                                  double f2, double4* __restrict__ srt, uint4* __restrict__ srf,
                                                  unsigned long long* __restrict__ dmax) {
    const int a = blockIdx.x * 64 + (threadIdx.x & 63), b = blockIdx.y * 4 + (threadIdx.x >> 6);
    __shared__ double red[4];
    double m = 0.0;
    if (a < na && b < nb) {
        const int m0 = a - a % 3;
        const double alpha2 = 0.5 * (1.0 - alpha);
        const double sc[3] = {s0, s1, s2}, lat[3] = {l0, l1, l2};
        const double rr[3] = {cb[a], cb[(size_t)na + a], cb[(size_t)2 * na + a]}, sr = cb[(size_t)3 * na + a];
        // the molecule's three atoms in every dimension, loaded before any is used
        double pm[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int u = 0; u < 3; u++) pm[d][u] = pos[((size_t)b * 3 + d) * na + m0 + u];
        double x[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = ((size_t)b * 3 + d) * na + a;
            if (a == m0) {
                const double v = alpha * pm[d][0] + alpha2 * (pm[d][1] + pm[d][2]);
                x[d] = sc[d] * v;
            } else {
                x[d] = sc[d] * (a == m0 + 1 ? pm[d][1] : pm[d][2]);
            }
            site[o] = x[d];
            if (ph) {
                double sn, cs;
                sincos(lat[d] * x[d], &sn, &cs);
                ph[o] = make_double2(cs, sn);
            }
        }
        const double dx = x[0] - rr[0], dy = x[1] - rr[1], dz = x[2] - rr[2];
        const double ex = fmax(sqrt(dx * dx + dy * dy + dz * dz) - sr, 0.0);
        double dd = ex * ex;
        if (!(dd <= 1e300)) dd = 1e300;
        m = dd;
        if (gather) {
            const size_t k = (size_t)b * (na + 1) + iperm[a];
            const double qq = q[a];
            srt[k] = make_double4(x[0], x[1], x[2], qq);
            srf[k] = site_fixed(x[0], x[1], x[2], qq, f0, f1, f2);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        ctrl_max(dmax, (unsigned long long)__double_as_longlong(fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))));
}

// When the list build runs (gate): per site its bead centroid c and spread s = max_b |x_b - c| from
// the sites of every bead, cu = [c_x, c_y, c_z, s][na].  Block: 64 sites x 4 bead lanes, lane g
// summing beads g, g + 4, ...; the lanes' sums are added in lane order.
__global__ __launch_bounds__(256) void k_vl_cu(const double* __restrict__ site, int na, int nb,
                                               double* __restrict__ cu, VGate gate) {
    if (!gate.on()) return;
    const int il = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int a = blockIdx.x * 64 + il;
    const bool av = a < na;
    __shared__ double part[4][64][3];
    __shared__ double red[4][64];
    double c[3] = {0.0, 0.0, 0.0};
    if (av)
        for (int b = g; b < nb; b += 4)
#pragma unroll
            for (int d = 0; d < 3; d++) c[d] += site[((size_t)b * 3 + d) * na + a];
#pragma unroll
    for (int d = 0; d < 3; d++) part[g][il][d] = c[d];
    __syncthreads();
    double cen[3];
#pragma unroll
    for (int d = 0; d < 3; d++) cen[d] = (((part[0][il][d] + part[1][il][d]) + part[2][il][d]) + part[3][il][d]) / nb;
    double s2m = 0.0;
    if (av)
        for (int b = g; b < nb; b += 4) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double e = site[((size_t)b * 3 + d) * na + a] - cen[d];
                r2 += e * e;
            }
            s2m = fmax(s2m, r2);
        }
    red[g][il] = s2m;
    __syncthreads();
    if (g == 0 && av) {
#pragma unroll
        for (int d = 0; d < 3; d++) cu[(size_t)d * na + a] = cen[d];
        cu[(size_t)3 * na + a] = sqrt(fmax(fmax(red[0][il], red[1][il]), fmax(red[2][il], red[3][il])));
    }
}

// One block, when the gate is on: a counting sort of the site centroids (cu, from k_vl_cu) into
// Morton cells, each cell then ordered by site index (perm / iperm, shared by all beads); the sorted
// centroids as wrapped fractional coordinates with the spreads (cs); the bounding box of each run of
// CLU sorted centroids (a cluster) with its largest spread, bb[c*8] = {centre, half-width, spread};
// and per cluster I the clusters J whose boxes come closer than rl + spread_I + spread_J (minimum
// image of the centres; L = box side, 1 for the unit box), ascending.  count must be all zero on entry
// and is left so.
__global__ __launch_bounds__(1024) void k_vl_sort(int na, CellGrid g, int nbins, double i0, double i1, double i2,
                                                  double L0, double L1, double L2, double rl, int nclu,
                                                  int* __restrict__ cell_of, int* __restrict__ count,
                                                  int* __restrict__ start, int* __restrict__ perm,
                                                  int* __restrict__ iperm, const double* __restrict__ cu,
                                                  float4* __restrict__ cs, double* __restrict__ bb,
                                                  int* __restrict__ ccnt, int* __restrict__ clist, VGate gate,
                                                  int* __restrict__ nreb) {
    if (!gate.on()) return;
    const int t = threadIdx.x;
    constexpr int NT = 1024;
    int* cn = count;
    int* st = start;
    // the sites are sorted again on every VL_RESORT-th build (and a forced one); between, the last
    // order is kept (any order gives the same pair set; the sort only keeps neighbours close)
    __shared__ int resort;
    if (t == 0) {
        const int k = ctrl_load(nreb);
        resort = gate.force || k % VL_RESORT == 0;
        ctrl_store(nreb, k + 1);
    }
    __syncthreads();
    if (!resort) {
        for (int x = t; x < na; x += NT) {
            const int a = perm[x];
            cs[x] = make_float4((float)wrap01(cu[a] * i0), (float)wrap01(cu[na + a] * i1),
                                (float)wrap01(cu[2 * na + a] * i2), (float)cu[3 * na + a]);
        }
    }
    for (int a = t; a < na && resort; a += NT) {
        const int cx = cell_coord(cu[a], i0, g.nc[0]);
        const int cy = cell_coord(cu[na + a], i1, g.nc[1]);
        const int cz = cell_coord(cu[2 * na + a], i2, g.nc[2]);
        const int cid = (int)(part1by2(cx) | (part1by2(cy) << 1) | (part1by2(cz) << 2));
        cell_of[a] = cid;
        atomicAdd(&cn[cid], 1);
    }
    __syncthreads();
    // exclusive scan of the counts into start; the counts restart from zero as fill counters
    __shared__ int buf[NT];
    __shared__ int carry;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nbins && resort; base += NT) {
        const int i = base + t;
        const int v = i < nbins ? cn[i] : 0;
        buf[t] = v;
        __syncthreads();
        for (int off = 1; off < NT; off <<= 1) {
            const int u = t >= off ? buf[t - off] : 0;
            __syncthreads();
            buf[t] += u;
            __syncthreads();
        }
        if (i < nbins) {
            st[i] = carry + buf[t] - v;
            cn[i] = 0;
        }
        __syncthreads();
        if (t == NT - 1) carry += buf[NT - 1];
        __syncthreads();
    }
    if (t == 0 && resort) st[nbins] = carry;
    __syncthreads();
    for (int a = t; a < na && resort; a += NT) {
        const int cid = cell_of[a];
        perm[st[cid] + atomicAdd(&cn[cid], 1)] = a;
    }
    __syncthreads();
    // order each cell by site index; sorted wrapped centroids and spreads
    for (int c = t; c < nbins && resort; c += NT) {
        cn[c] = 0;
        const int s0 = st[c], s1 = st[c + 1];
        for (int x = s0 + 1; x < s1; x++) {
            const int v = perm[x];
            int y = x - 1;
            while (y >= s0 && perm[y] > v) {
                perm[y + 1] = perm[y];
                y--;
            }
            perm[y + 1] = v;
        }
        for (int x = s0; x < s1; x++) {
            const int a = perm[x];
            cs[x] = make_float4((float)wrap01(cu[a] * i0), (float)wrap01(cu[na + a] * i1),
                                (float)wrap01(cu[2 * na + a] * i2), (float)cu[3 * na + a]);
            iperm[a] = x;
        }
    }
    __syncthreads();
    // cluster boxes (FP64 from the FP32 wrapped centroids), wave w taking clusters w, w + NT/64, ...
    const int w = t >> 6, l = t & 63;
    constexpr int NW = NT / 64;
    for (int c = w; c < nclu; c += NW) {
        const int a = c * CLU + l;
        double lo[3], hi[3], sp = 0.0;
        const float4 v = a < na ? cs[a] : make_float4(0.f, 0.f, 0.f, 0.f);
        const double f[3] = {v.x, v.y, v.z};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = a < na ? f[d] : 2.0;
            hi[d] = a < na ? f[d] : -1.0;
        }
        sp = a < na ? (double)v.w : 0.0;
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
                hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
            }
            sp = fmax(sp, __shfl_xor(sp, off, 64));
        }
        if (l == 0) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bb[c * 8 + d] = 0.5 * (lo[d] + hi[d]);
                bb[c * 8 + 3 + d] = 0.5 * (hi[d] - lo[d]);
            }
            bb[c * 8 + 6] = sp;
        }
    }
    __syncthreads();
    const double L[3] = {L0, L1, L2};
    const double lmax = fmax(L0, fmax(L1, L2));
    for (int I = w; I < nclu; I += NW) {
        const double* bi = bb + I * 8;
        int* out = clist + (size_t)I * nclu;
        int n = 0;
        for (int J0 = 0; J0 < nclu; J0 += 64) {
            const int J = J0 + l;
            bool keep = false;
            if (J < nclu) {
                const double* bj = bb + J * 8;
                double r2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double dc = bi[d] - bj[d];
                    dc -= rint(dc);
                    const double gap = fmax(0.0, fabs(dc) - bi[3 + d] - bj[3 + d]) * L[d];
                    r2 += gap * gap;
                }
                // reach with a margin for the FP32 centroids (1e-5 of the largest box side)
                const double reach = rl + bi[6] + bj[6] + 1e-5 * lmax;
                keep = r2 < reach * reach;
            }
            const uint64_t m = __ballot(keep);
            if (keep) out[n + __popcll(m & ((1ull << l) - 1))] = J;
            n += __popcll(m);
        }
        if (l == 0) ccnt[I] = n;
    }
}

// Per sorted site k its list of sorted neighbours j with |c_k - c_j| < rl + s_k + s_j (centroids c,
// spreads s): every bead pair then lies within rl of each other only if listed, because
// |x_kb - x_jb| >= |c_k - c_j| - s_k - s_j.  First the other sites of its own molecule (nex[] of them;
// the real-space term treats those pairs as excluded, src/ewald_cubic.f90:134), then the others in
// ascending sorted order, each entry stored as 16 j (the byte offset of site j's 16-byte FP32 coordinate
// record).  Entry n of site k is nl[((n/VLG)*na + k)*VLG + (n%VLS)*4 + (n/VLS)%4]: lane
// n % VLS of the VLS lanes that walk a site's list reads its next four entries as one 16-byte load,
// the site's lanes read 64 contiguous bytes, and consecutive entries are usually neighbouring sites.
// The last group of a site is filled up with the pad site na (no charge), and group cap/VLG holds it
// sixteen times (the pair kernel reads it past a lane's end).  The test is
// FP32 on wrapped fractional centroids against the reach enlarged by a margin that covers the FP32
// error, so the list is a superset (a few pairs beyond do no harm: the pair kernel tests the
// cut-off exactly).  One wave per sorted site k (four per block), a lane per candidate: for each listed
// cluster J (ascending) lane l tests site 64 J + l, the candidates of the next cluster loaded while the
// current ones are tested, and the hits are stored at their ascending positions from the wave's mask.
// The block also records the centroid and spread of its sites (the reference of the displacement
// gate) and writes their current coordinates in the new order, as k_sites_vl does between builds.
constexpr int BLS = 4;
static_assert(CLU % BLS == 0, "a block's sites lie in one cluster");
__global__ __launch_bounds__(64 * BLS) void k_vl_build(const float4* __restrict__ cs, const int* __restrict__ perm,
                                                       const int* __restrict__ iperm, const double* __restrict__ site,
                                                       const double* __restrict__ q, int na, int nb, int nclu,
                                                       const int* __restrict__ ccnt, const int* __restrict__ clist,
                                                       float L0, float L1, float L2, double f0, double f1, double f2,
                                                       float rl, float mrg, int cap, int* __restrict__ cnt,
                                                       int* __restrict__ nex, int* __restrict__ nl,
                                                       const double* __restrict__ cu, double* __restrict__ cb,
                                                       double4* __restrict__ srt, uint4* __restrict__ srf,
                                                       int* __restrict__ ovfl, VGate gate) {
    if (!gate.on()) return;
    const int t = threadIdx.x, l = t & 63;
    const int k0 = blockIdx.x * BLS, k = k0 + (t >> 6);
    // the build's centroid and spread of the block's sites (the displacement gate's reference)
    if (t < 4 * BLS && k0 + t % BLS < na) {
        const int a = perm[k0 + t % BLS];
        cb[(size_t)(t / BLS) * na + a] = cu[(size_t)(t / BLS) * na + a];
    }
    // sorted current coordinates of the block's sites, every bead
    for (int e = t; e < BLS * nb; e += 64 * BLS) {
        const int kk = k0 + e % BLS, b = e / BLS;
        if (kk < na) {
            const int a = perm[kk];
            const size_t o = (size_t)b * 3 * na + a;
            const double x = site[o], y = site[o + na], z = site[o + 2 * na], qq = q[a];
            srt[(size_t)b * (na + 1) + kk] = make_double4(x, y, z, qq);
            srf[(size_t)b * (na + 1) + kk] = site_fixed(x, y, z, qq, f0, f1, f2);
        }
    }
    // the charge-free pad site of every bead
    if (blockIdx.x == 0)
        for (int b = t; b < nb; b += 64 * BLS) {
            srt[(size_t)b * (na + 1) + na] = make_double4(0.0, 0.0, 0.0, 0.0);
            srf[(size_t)b * (na + 1) + na] = make_uint4(1u, 1u, 1u, 0u);
        }
    if (k >= na) return;  // a whole wave (k is the wave's)
    const float4 ci = cs[k];
    const int a = perm[k], mi = a / 3, I = k / CLU;
    int n = 0;
    bool ovf = false;
    auto put = [&](int p, int j) {
        if (p < cap) nl[((size_t)(p / VLG) * na + k) * VLG + (p % VLS) * 4 + (p / VLS) % 4] = 16 * j;
        else ovf = true;
    };
    auto reach2 = [&](float sj) {
        const float r = rl + ci.w + sj;
        return r * r * mrg;
    };
    auto d2 = [&](float4 c) {
        float dx = ci.x - c.x, dy = ci.y - c.y, dz = ci.z - c.z;
        dx = (dx - rintf(dx)) * L0;
        dy = (dy - rintf(dy)) * L1;
        dz = (dz - rintf(dz)) * L2;
        return dx * dx + dy * dy + dz * dz;
    };
    if (l == 0) {
        const int m0 = 3 * mi;
        for (int p = m0; p < m0 + 3; p++) {
            if (p == a || p >= na) continue;
            const int jk = iperm[p];
            const float4 c = cs[jk];
            if (d2(c) < reach2(c.w)) put(n++, jk);
        }
        nex[k] = min(n, cap);
    }
    n = __shfl(n, 0, 64);  // the own-molecule entries, from lane 0
    const int nJ = ccnt[I];
    const int* lst = clist + (size_t)I * nclu;
    // lane l's candidate of listed cluster e: its centroid and spread, and its molecule (-1: none)
    auto cand = [&](int e, float4& c, int& m) {
        const int j = lst[e] * CLU + l;
        c = j < na ? cs[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        m = j < na ? perm[j] / 3 : -1;
    };
    const uint64_t below = (1ull << l) - 1ull;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    int mj = -1;
    if (nJ > 0) cand(0, c, mj);
    for (int e = 0; e < nJ; e++) {
        const int J = lst[e];
        float4 cn = c;
        int mn = -1;
        if (e + 1 < nJ) cand(e + 1, cn, mn);
        const bool take = mj >= 0 && mj != mi && d2(c) < reach2(c.w);
        const uint64_t m = __ballot(take);
        if (take) put(n + __popcll(m & below), J * CLU + l);
        n += __popcll(m);
        c = cn;
        mj = mn;
    }
    // the rest of the site's last group and the padding group hold the pad site na
    const int nf = min(n, cap), pe = min((nf + VLG - 1) / VLG * VLG, cap);
    for (int p = nf + l; p < pe; p += 64) put(p, na);
    if (l < VLG) nl[((size_t)(cap / VLG) * na + k) * VLG + l] = 16 * na;
    if (l == 0) cnt[k] = nf;
    if (ovf) ovfl[0] = 1;
}

// one pair of the list walk: gradient and the 7 independent sums {v, xx, xy, xz, yy, yz, zz}
// (the pair force is central, so the virial dx_a f_b is symmetric)
template <typename T>
__device__ __forceinline__ void vl_pair(T dx, T dy, T dz, T dr2, T qij, bool same, T ta, T tsf, T (&tg)[3],
                                        T (&tv)[7]) {
    T du, dur;
    erfc_pair<T>(dr2, ta, tsf, same, du, dur);
    const T dv = qij * du, dvr = qij * dur;
    const T fxx = dvr * dx, fyy = dvr * dy, fzz = dvr * dz;
    tg[0] += fxx;
    tg[1] += fyy;
    tg[2] += fzz;
    tv[0] += dv;
    tv[1] += dx * fxx;
    tv[2] += dx * fyy;
    tv[3] += dx * fzz;
    tv[4] += dy * fyy;
    tv[5] += dy * fzz;
    tv[6] += dz * fzz;
}

// FP32 pair sums: gradient and the 7 sums as vl_pair, from du and dur, without the site's own charge
// (a factor of all of a site's sums, applied when they are flushed)
__device__ __forceinline__ void pair_add(float dx, float dy, float dz, float qj, float du, float dur, float (&tg)[3],
                                         float (&tv)[7]) {
    const float dvr = qj * dur;
    const float fxx = dvr * dx, fyy = dvr * dy, fzz = dvr * dz;
    tg[0] += fxx;
    tg[1] += fyy;
    tg[2] += fzz;
    tv[0] += qj * du;
    tv[1] += dx * fxx;
    tv[2] += dx * fyy;
    tv[3] += dx * fzz;
    tv[4] += dy * fyy;
    tv[5] += dy * fzz;
    tv[6] += dz * fzz;
}
// FP32 pair body on fixed-point coordinates: du and dur (erfc, in the tables' scales 2^ku, 2^kr) from
// the pair-function table in LDS at the scaled abscissa sp = c r^2 (PairTab: clamped to the table, the
// interval is the exponent and top PT_M mantissa bits, the abscissa w the remaining mantissa bits,
// exact in FP32; at and past the cut-off the zero interval), no transcendental and no cut-off test.
// Own-molecule pairs (erfc - 1) and pairs below the table are corrected after the walk
// (ewald_vl_body).
__device__ __forceinline__ void vl_pair_tab(float dx, float dy, float dz, float sp, float qj,
                                            const float4* __restrict__ tl, float (&tg)[3], float (&tv)[7]) {
    const uint32_t bits = __float_as_uint(__builtin_amdgcn_fmed3f(sp, 2.0f, 0x1p14f));
    const float w = (float)(bits & ((1u << (23 - PT_M)) - 1u));
    // byte offset of the interval's entries (16 bytes each, du's then dur's array): the low 4
    // exponent bits and the top PT_M mantissa bits, by a mask (known non-negative, so the arrays'
    // bases go into the LDS instructions' offsets); the 16-byte stride spreads a wave's reads over
    // all LDS banks
    const uint32_t off = (bits >> (23 - PT_M - 4)) & (((16u << PT_M) - 1u) << 4);
    const char* tb = (const char*)tl;
    const float4 A = *(const float4*)(tb + off), B = *(const float4*)(tb + PT_N * 16 + off);
    const float du = fmaf(fmaf(fmaf(A.w, w, A.z), w, A.y), w, A.x);
    const float dur = fmaf(fmaf(fmaf(B.w, w, B.z), w, B.y), w, B.x);
    pair_add(dx, dy, dz, qj, du, dur, tg, tv);
}
// the A&S form itself in FP32 (erfc, or erfc - 1 for an own-molecule pair), in the tables' scales;
// ta = alpha, tsf = sfac in the kernel's distance units
__device__ __forceinline__ void erfc_pair_scaled(float r2, float ta, float tsf, bool same, float cu, float cr,
                                                 float& du, float& dur) {
    const float ri = __frsqrt_rn(r2);
    const float x = ta * (r2 * ri);
    const float ex = __expf(-x * x);
    const float t = __builtin_amdgcn_rcpf((float)AS_P + x) * (float)AS_P;
    float h = ex * t * ((float)AS_A1 + t * ((float)AS_A2 + t * ((float)AS_A3 + t * ((float)AS_A4 + t * (float)AS_A5))));
    if (same) h -= 1.0f;
    const float d = h * ri;
    du = d * cu;
    dur = ((tsf * ex - d) * (ri * ri)) * cr;
}

// sum over the four lanes of a quad (lanes 4i..4i+3), every lane gets the same bits
__device__ __forceinline__ float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // lanes ^ 1
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // lanes ^ 2
    return v;
}
__device__ __forceinline__ double quad_sum(double v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}

// VLS lanes per sorted site, lane s walking entries s, s + VLS, ... of the site's list (one list for
// all beads).  The walk takes the lane's entries in groups of four with two register sets: the
// coordinates of one group and the list entries of the group after it are in flight while a group is
// computed (no copy of a loaded register, which would wait for its load).  Entries past a site's list
// are the charge-free pad site na (FP64: also masked by position).  The site's gradient is summed over
// its lanes by a fixed butterfly.  Block id i runs site block (i % 8) + 8 ((i / 8) / nb) of bead
// (i / 8) % nb, so the nb blocks that read one stretch of the list run on one XCD (blocks go to the 8
// XCDs round robin) and share its L2.  Clears dmax for the next call (every gated kernel has run).
// FP32: the differences of the fixed-point coordinates (srf) are the minimum image, converted to FP32
// exactly up to rounding; for the unit box they stay in units of 2^-32 (the constants and the final
// sums are scaled by powers of two instead), else they are scaled by L 2^-32 per component.  The pad
// site has odd coordinates and the others even ones, so no distance is zero, and its pairs add zero.
#define EWVL_PARAMS \
    const double4* __restrict__ srt, const uint4* __restrict__ srf, const int* __restrict__ perm, \
    const int* __restrict__ cnt, const int* __restrict__ nex, const int* __restrict__ nl, int na, int nb, int cap, \
    double L0, double L1, double L2, double i0, double i1, double i2, double rc2, double alpha, double sfac, \
    double* __restrict__ grad, double* __restrict__ parts, unsigned long long* __restrict__ dmax, \
    const float4* __restrict__ tab, float tsc, int tku, int tkr
#define EWVL_ARGS                                                                                                  \
    srt, srf, perm, cnt, nex, nl, na, nb, cap, L0, L1, L2, i0, i1, i2, rc2, alpha, sfac, grad, parts, dmax, tab, tsc, \
        tku, tkr
template <typename T, bool UNIT>
__device__ __forceinline__ void ewald_vl_body(EWVL_PARAMS, const int id) {
    constexpr bool F64 = sizeof(T) == 8;
    typedef typename std::conditional<F64, double4, uint4>::type P4;
    const int m8 = id >> 3;
    const int b = m8 % nb, sb = (id & 7) + 8 * (m8 / nb);
    const int k = (sb * blockDim.x + threadIdx.x) / VLS, sl = threadIdx.x % VLS;
    if (id == 0 && threadIdx.x == 0) ctrl_store(dmax, 0ull);
    // LDS: the FP64 sums of each site (energy, virial, gradient; fewer registers, more waves), added
    // to by the site's first lane, then (FP32) the pair-function table, copied in by the block, at
    // a fixed address that the LDS instructions take as offset.
    struct Lds {
        double vacc[10][256 / VLS];
        float qs[256 / VLS];  // FP32: the site's charge, for the end (no register held through the walk)
        float4 tl[F64 ? 1 : 2 * PT_N];
    };
    __shared__ Lds lds;
    double(&vacc)[10][256 / VLS] = lds.vacc;
    const float4* tl = lds.tl;
    if (sl == 0)
#pragma unroll
        for (int c = 0; c < 10; c++) vacc[c][threadIdx.x / VLS] = 0.0;
    if constexpr (!F64) {
        for (int i = threadIdx.x; i < 2 * PT_N; i += blockDim.x) lds.tl[i] = tab[i];
        __syncthreads();
    }
    // FP32 unit box: distances in units of 2^-32, gradient x 2^64, energy and virial x 2^32 at the end
    constexpr double SR = (!F64 && UNIT) ? 0x1p-32 : 1.0;
    // FP32: two register sets of coordinates (a group in flight); FP64 (twice the registers): one
    constexpr bool DB = !F64;
    if (k < na) {
        const size_t bo = (size_t)b * (na + 1);
        const P4* X;
        if constexpr (F64)
            X = srt + bo;
        else
            X = srf + bo;
        // this lane's groups of four entries: uint4 (g na + k) VLS + sl, addressed by 32-bit byte
        // offsets from the uniform list base (the list is far below 4 GB)
        const uint32_t l0 = (uint32_t)(k * VLS + sl) * 16u, ls = (uint32_t)na * VLS * 16u;
        const int ntot = cnt[k], ne = nex[k];
        const int nn = (ntot - sl + VLS - 1) / VLS;  // entries of this lane
        // groups: those of the site's first lane, for all its lanes (past a lane's entries the groups
        // hold the pad site), so the site's lanes stay together for the sums over a quad
        const int ng = ((ntot + VLS - 1) / VLS + 3) / 4;
        const P4 xi = X[k];
        T qi;
        if constexpr (F64) {
            qi = xi.w;
        } else {
            qi = __uint_as_float(xi.w);
            if (sl == 0) lds.qs[threadIdx.x / VLS] = qi;
        }
        const T Lt0 = (T)L0, Lt1 = (T)L1, Lt2 = (T)L2, it0 = (T)i0, it1 = (T)i1, it2 = (T)i2;
        const float Ls0 = (float)(L0 * 0x1p-32), Ls1 = (float)(L1 * 0x1p-32), Ls2 = (float)(L2 * 0x1p-32);
        const T trc2 = (T)(rc2 / (SR * SR));
        const T ta = (T)(alpha * SR), tsf = (T)(sfac * SR);
        const float cu = __builtin_ldexpf(1.0f, tku), cr = __builtin_ldexpf(1.0f, tkr);
        T tg[3] = {0, 0, 0}, tv[7];
#pragma unroll
        for (int c = 0; c < 7; c++) tv[c] = 0;
        float lo = 3.0e38f;  // FP32: the lane's smallest scaled r^2
        // past the lane's last group: group cap / VLG, which holds the pad site (no branch around the
        // load)
        const int pad = cap / VLG;
        auto ent = [&](int g) {
            return *(const uint4*)((const char*)nl + (l0 + (uint32_t)(g < ng ? g : pad) * ls));
        };
        // coordinates of the site of list entry e (the byte offset of the site's 16-byte FP32 record,
        // half that of its FP64 one), by a 32-bit byte offset from the bead's base
        auto at = [&](uint32_t e) { return *(const P4*)((const char*)X + e * (uint32_t)(sizeof(P4) / 16)); };
        // a site has at most two own-molecule partners (three sites per molecule), listed first, so
        // only a lane's first entry can be one: may_same is false for every other call
        // FP32: coordinate difference (as float in the kernel's units), r^2, the table's abscissa
        // c r^2 and the charge of the site with list entry cur
        auto diff32 = [&](const P4& cur, float& ex, float& ey, float& ez, float& r2, float& sp, float& qj) {
            ex = (float)(int)(xi.x - cur.x);
            ey = (float)(int)(xi.y - cur.y);
            ez = (float)(int)(xi.z - cur.z);
            if (!UNIT) {
                ex *= Ls0;
                ey *= Ls1;
                ez *= Ls2;
            }
            r2 = ex * ex + ey * ey + ez * ez;
            sp = tsc * r2;
            qj = __uint_as_float(cur.w);
        };
        auto pair = [&](const P4& cur, int r, bool may_same) {
            if constexpr (F64) {
                T ex = xi.x - cur.x, ey = xi.y - cur.y, ez = xi.z - cur.z;
                if (UNIT) {
                    // the unit box (cubic): both forms reduce to this exactly
                    ex -= rint(ex);
                    ey -= rint(ey);
                    ez -= rint(ez);
                } else {
                    // physical minimum image
                    ex -= Lt0 * rint(ex * it0);
                    ey -= Lt1 * rint(ey * it1);
                    ez -= Lt2 * rint(ez * it2);
                }
                const T r2 = ex * ex + ey * ey + ez * ez;
                const bool in = r < nn && r2 < trc2;
                vl_pair<T>(ex, ey, ez, in ? r2 : (T)1, in ? qi * (T)cur.w : (T)0, may_same && r * VLS + sl < ne,
                           ta, tsf, tg, tv);
            } else {
                float ex, ey, ez, r2, sp, qj;
                diff32(cur, ex, ey, ez, r2, sp, qj);
                lo = fminf(lo, sp);
                vl_pair_tab(ex, ey, ez, sp, qj, tl, tg, tv);
            }
        };
        // the lanes' sums of a site in a fixed order (a quad's butterfly), added to its FP64 sums
        auto flush = [&]() {
#pragma unroll
            for (int c = 0; c < 10; c++) {
                const T t = quad_sum(c < 7 ? tv[c] : tg[c - 7]);
                if (sl == 0) vacc[c][threadIdx.x / VLS] += (double)t;
            }
#pragma unroll
            for (int c = 0; c < 7; c++) tv[c] = 0;
            tg[0] = tg[1] = tg[2] = 0;
        };
        // groups g and g + 1 per trip, also when g + 1 is past the lane's end (the pad group then, whose
        // pairs add zero): with no exit between a load and its use the loads stay where they are
        // issued, a group ahead of their use
        if constexpr (DB) {
            uint4 ia = ent(0);
            P4 a0 = at(ia.x), a1 = at(ia.y), a2 = at(ia.z), a3 = at(ia.w);
            uint4 ib = ent(1);
            for (int g = 0; g < ng; g += 2) {
                const int r = 4 * g;
                const P4 b0 = at(ib.x), b1 = at(ib.y), b2 = at(ib.z), b3 = at(ib.w);
                ia = ent(g + 2);
                pair(a0, r, g == 0);
                pair(a1, r + 1, false);
                pair(a2, r + 2, false);
                pair(a3, r + 3, false);
                a0 = at(ia.x);
                a1 = at(ia.y);
                a2 = at(ia.z);
                a3 = at(ia.w);
                ib = ent(g + 3);
                pair(b0, r + 4, false);
                pair(b1, r + 5, false);
                pair(b2, r + 6, false);
                pair(b3, r + 7, false);
                if (!F64 && (g & 15) == 14) flush();
            }
        } else {
            uint4 ia = ent(0);
            for (int g = 0; g < ng; g++) {
                const int r = 4 * g;
                const P4 a0 = at(ia.x), a1 = at(ia.y), a2 = at(ia.z), a3 = at(ia.w);
                ia = ent(g + 1);
                pair(a0, r, g == 0);
                pair(a1, r + 1, false);
                pair(a2, r + 2, false);
                pair(a3, r + 3, false);
                if (!F64 && (g & 3) == 3) flush();
            }
        }
        if constexpr (!F64) {
            // the lane's first entry, when it is of the site's own molecule: the -1 of erfc - 1
            // (du - 1/r, dur + 1/r^3)
            if (sl < ne) {
                float ex, ey, ez, r2, sp, qj;
                diff32(at(ent(0).x), ex, ey, ez, r2, sp, qj);
                const float ri = __frsqrt_rn(r2);
                pair_add(ex, ey, ez, sp < 0x1p14f ? qj : 0.0f, -cu * ri, cr * (ri * ri) * ri, tg, tv);
            }
            // pairs closer than the table's first node (no physical configuration has them): the A&S
            // form in place of the table's value there, which the walk added
            if (lo < 2.0f) {
                const float du0 = tl[0].x, dur0 = tl[PT_N].x;
                for (int g = 0; g < ng; g++) {
                    const uint4 ie = ent(g);
                    const uint32_t js[4] = {ie.x, ie.y, ie.z, ie.w};
                    for (int q = 0; q < 4; q++) {
                        float ex, ey, ez, r2, sp, qj, du, dur;
                        diff32(at(js[q]), ex, ey, ez, r2, sp, qj);
                        if (sp < 2.0f) {
                            erfc_pair_scaled(r2, ta, tsf, false, cu, cr, du, dur);
                            pair_add(ex, ey, ez, qj, du - du0, dur - dur0, tg, tv);
                        }
                    }
                }
            }
        }
        flush();
    }
    // FP32: gradient and virial in the dur table's scale, energy in the du table's, all without the
    // site's charge (kept in LDS through the walk)
    const bool own = k < na && sl == 0;
    const double qs = F64 || !own ? 1.0 : (double)lds.qs[threadIdx.x / VLS];
    const double GS = F64 ? 1.0 : ldexp(1.0 / (SR * SR), -tkr) * qs;
    double acc[7];
    constexpr double VS = 0.5 / SR;
    const double VU = F64 ? VS : ldexp(VS, -tku) * qs, VR = F64 ? VS : ldexp(VS, -tkr) * qs;
#pragma unroll
    for (int c = 0; c < 7; c++) acc[c] = own ? vacc[c][threadIdx.x / VLS] * (c == 0 ? VU : VR) : 0.0;
    if (own) {
        const int a = perm[k], sk = threadIdx.x / VLS;
        double* G = grad + (size_t)b * 3 * na;
        G[a] = vacc[7][sk] * GS;
        G[na + a] = vacc[8][sk] * GS;
        G[2 * na + a] = vacc[9][sk] * GS;
    }
    // {v, vir(3,3) column-major} from the symmetric sums
    double w[10] = {acc[0], acc[1], acc[2], acc[3], acc[2], acc[4], acc[5], acc[3], acc[5], acc[6]};
    block_reduce_store<10>(w, parts + ((size_t)b * ewald_real_parts(na) + sb) * 10);
}

// FP32: held to 64 VGPRs, so that 8 waves of it fit a SIMD; FP64 (twice the registers per value)
// as the compiler chooses
template <bool UNIT>
__global__ __launch_bounds__(256, 8) void k_ewald_vl32(EWVL_PARAMS) {
    ewald_vl_body<float, UNIT>(EWVL_ARGS, blockIdx.x);
}
template <bool UNIT>
__global__ __launch_bounds__(256) void k_ewald_vl64(EWVL_PARAMS) {
    ewald_vl_body<double, UNIT>(EWVL_ARGS, blockIdx.x);
}

int ewald_nclu(int na) { return (na + CLU - 1) / CLU; }
int ewald_real_slots() { return 1; }

// FP32 pair-test margin (relative, on r^2): wrapped fractional FP32 coordinates carry an absolute
// error of about 2^-24 per component, so |dr| errs by less than 1e-6 of the box side; the margin
// 1e-4 + 1e-5 L_max / r covers that for any cut-off r
static double pair_margin(const Box& sb, double r) {
    const double lmax = fmax(sb.L[0], fmax(sb.L[1], sb.L[2]));
    return 1e-4 + 1e-5 * lmax / r;
}

// expected neighbours within rl of a site, with margin, as the list capacity
int vlist_cap(int na, Box sb, double rl) {
    const double vol = sb.L[0] * sb.L[1] * sb.L[2];
    const double frac = 4.0 / 3.0 * acos(-1.0) * rl * rl * rl / vol;
    double c = 1.5 * frac * na + 64.0;
    if (c > na - 1) c = na - 1;
    int cap = (int)ceil(c);
    if (cap < 1) cap = 1;
    return (cap + VLG - 1) / VLG * VLG;
}

void launch_sites_vl(const double* pos, double* site, int na, int nb, double alpha, int unit, Box box,
                     const double lat[3], double* ph, const double* q, const VList& v, int gather, Box sb,
                     hipStream_t s) {
    const double s0 = unit ? box.inv[0] : 1.0, s1 = unit ? box.inv[1] : 1.0, s2 = unit ? box.inv[2] : 1.0;
    k_sites_vl<<<dim3((unsigned)((na + 63) / 64), (unsigned)((nb + 3) / 4)), 256, 0, s>>>(
        pos, site, na, nb, alpha, s0, s1, s2, lat[0], lat[1], lat[2], (double2*)ph, q, v.cb, v.iperm, gather,
        sb.inv[0], sb.inv[1], sb.inv[2], v.srt, v.srf, v.dmax);
}

// The phase factors alone (the same values as k_sites_vl writes), thread per (site, bead): the k-space
// branch computes them itself when it runs beside the real-space branch
__global__ __launch_bounds__(256) void k_phases(const double* __restrict__ pos, int na, int nb, double alpha,
                                                double s0, double s1, double s2, double l0, double l1, double l2,
                                                double2* __restrict__ ph) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)na * nb) return;
    const int b = (int)(i / na), a = (int)(i % na);
    const int m0 = a - a % 3;
    const double alpha2 = 0.5 * (1.0 - alpha);
    const double sc[3] = {s0, s1, s2}, lat[3] = {l0, l1, l2};
    // the molecule's three atoms in every dimension, loaded before any is used
    double pm[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int u = 0; u < 3; u++) pm[d][u] = pos[((size_t)b * 3 + d) * na + m0 + u];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double pa = a == m0 ? pm[d][0] : (a == m0 + 1 ? pm[d][1] : pm[d][2]);
        double x;
        if (a == m0) {
            const double v = alpha * pm[d][0] + alpha2 * (pm[d][1] + pm[d][2]);
            x = sc[d] * v;
        } else {
            x = sc[d] * pa;
        }
        double sn, cs;
        sincos(lat[d] * x, &sn, &cs);
        ph[((size_t)b * 3 + d) * na + a] = make_double2(cs, sn);
    }
}
void launch_phases(const double* pos, int na, int nb, double alpha, int unit, Box box, const double lat[3], double* ph,
                   hipStream_t s) {
    const double s0 = unit ? box.inv[0] : 1.0, s1 = unit ? box.inv[1] : 1.0, s2 = unit ? box.inv[2] : 1.0;
    const size_t n = (size_t)na * nb;
    k_phases<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pos, na, nb, alpha, s0, s1, s2, lat[0], lat[1], lat[2],
                                                         (double2*)ph);
}

// The build runs when a bead site has left its build sphere (centroid, spread) by more than skin/2
// (or force is set); it lists, for all beads at once, the pairs whose centroids are within
// rl + s_k + s_j, rl = rc + skin.  count: [nbins + 1] ints, all zero.
void launch_vlist(const VList& v, const double* site, const double* q, int na, int nb, Box sb, CellGrid g, double rc,
                  double skin, int force, int* cell_of, int* count, int* start, hipStream_t s) {
    const VGate gate{v.dmax, 0.25 * skin * skin, force};
    int p = 1;
    while (p < g.nc[0] || p < g.nc[1] || p < g.nc[2]) p <<= 1;
    const int nbins = p * p * p;
    const int nclu = ewald_nclu(na);
    const double rl = rc + skin;
    k_vl_cu<<<(unsigned)((na + 63) / 64), 256, 0, s>>>(site, na, nb, v.cu, gate);
    k_vl_sort<<<1, 1024, 0, s>>>(na, g, nbins, sb.inv[0], sb.inv[1], sb.inv[2], sb.L[0], sb.L[1], sb.L[2], rl, nclu,
                                 cell_of, count, start, v.perm, v.iperm, v.cu, v.cs, v.bb, v.ccnt, v.clist, gate,
                                 v.nreb);
    const float mrg = (float)(1.0 + pair_margin(sb, rl));
    k_vl_build<<<(na + BLS - 1) / BLS, 64 * BLS, 0, s>>>(v.cs, v.perm, v.iperm, site, q, na, nb, nclu, v.ccnt, v.clist,
                                                  (float)sb.L[0], (float)sb.L[1], (float)sb.L[2], sb.inv[0],
                                                  sb.inv[1], sb.inv[2], (float)rl, mrg, v.cap, v.cnt, v.nex, v.nl,
                                                  v.cu, v.cb, v.srt, v.srf, v.gate + 1, gate);
}

// Dynamic LDS for a kernel so that at most bpc of its blocks fit on a CU (0: none, or bpc <= 0).
static size_t lds_for_blocks_per_cu(const void* fn, int bpc) {
    if (bpc <= 0) return 0;
    int dev = 0, lds_cu = 0;
    hipFuncAttributes fa{};
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, fn) != hipSuccess || lds_cu <= 0)
        return 0;
    size_t total = (size_t)lds_cu / (bpc + 1) + 1024;
    if (total > (size_t)lds_cu / bpc) total = (size_t)lds_cu / bpc;
    return total > fa.sharedSizeBytes ? total - fa.sharedSizeBytes : 0;
}

// Pair-function table (PairTab): on interval i of octave o (s' from 2^e (1 + j/2^PT_M), width
// 2^(e - PT_M), e = 1 + o) each function is the cubic through its values at the four Chebyshev
// nodes of the interval, in w = (s' - s'_i) 2^(23 - e) (the low mantissa bits of s'), computed in FP64
// at s = s' / c from the A&S form of the reference (src/ewald_cubic.f90:126-140), then scaled by a
// power of two that brings the largest coefficient near 1 and rounded to FP32.  c is the FP32 value
// the kernel multiplies by.
size_t pair_table_bytes() { return (size_t)PT_N * 2 * sizeof(float4); }
void pair_table_fill(double alpha, double rc, double sr, PairTab& t, float4* h) {
    const int per = 1 << PT_M, n = PT_N - 1;
    const double a = alpha * sr, sfac = -2.0 * (alpha / sqrt(acos(-1.0))) * sr, s_rc = (rc / sr) * (rc / sr);
    t.c = (float)(ldexp(1.0, 1 + PT_OCT) / s_rc);
    const double ic = 1.0 / (double)t.c;
    auto du_dur = [&](double s, double& du, double& dur) {
        const double r = sqrt(s), x = a * r, e = exp(-x * x), tt = AS_P / (AS_P + x);
        du = e * tt * (AS_A1 + tt * (AS_A2 + tt * (AS_A3 + tt * (AS_A4 + tt * AS_A5)))) / r;
        dur = (sfac * e - du) / s;
    };
    // inverse of the Vandermonde matrix of the nodes v_k in [0, 1]
    double v[4], inv[4][4];
    for (int k = 0; k < 4; k++) v[k] = 0.5 * (1.0 - cos((2 * k + 1) * acos(-1.0) / 8.0));
    for (int i = 0; i < 4; i++) {
        // Lagrange basis polynomial i expanded in powers of v
        double c[4] = {1.0, 0.0, 0.0, 0.0}, den = 1.0;
        int deg = 0;
        for (int j = 0; j < 4; j++) {
            if (j == i) continue;
            for (int m = deg + 1; m > 0; m--) c[m] = c[m - 1] - v[j] * c[m];
            c[0] = -v[j] * c[0];
            deg++;
            den *= v[i] - v[j];
        }
        for (int m = 0; m < 4; m++) inv[m][i] = c[m] / den;
    }
    std::vector<double> cf((size_t)n * 8);
    const double W = ldexp(1.0, -(23 - PT_M));  // v per unit of w
    for (int i = 0; i < n; i++) {
        const int o = i / per, j = i % per;
        const double base = ldexp(1.0, 1 + o), s0 = base * (1.0 + (double)j / per), hw = base / per;
        double f[2][4];
        for (int k = 0; k < 4; k++) du_dur((s0 + v[k] * hw) * ic, f[0][k], f[1][k]);
        for (int q = 0; q < 2; q++)
            for (int m = 0; m < 4; m++) {
                double acc = 0.0;
                for (int k = 0; k < 4; k++) acc += inv[m][k] * f[q][k];
                cf[(size_t)i * 8 + q * 4 + m] = acc * pow(W, m);
            }
    }
    double mx[2] = {0.0, 0.0};
    for (int i = 0; i < n; i++)
        for (int q = 0; q < 2; q++) mx[q] = fmax(mx[q], fabs(cf[(size_t)i * 8 + q * 4]));
    t.ku = -ilogb(mx[0]);
    t.kr = -ilogb(mx[1]);
    for (int i = 0; i < n; i++) {
        const double* c = &cf[(size_t)i * 8];
        h[i] = make_float4((float)ldexp(c[0], t.ku), (float)ldexp(c[1], t.ku), (float)ldexp(c[2], t.ku),
                           (float)ldexp(c[3], t.ku));
        h[PT_N + i] = make_float4((float)ldexp(c[4], t.kr), (float)ldexp(c[5], t.kr), (float)ldexp(c[6], t.kr),
                                  (float)ldexp(c[7], t.kr));
    }
    h[n] = h[PT_N + n] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // at and past the cut-off
    t.key[0] = alpha;
    t.key[1] = rc;
    t.key[2] = sr;
}

// bpc > 0: at most bpc blocks per CU (FP32 pair kernel, which otherwise fills every CU slot), so that
// kernels on another stream keep CU slots beside it
template <typename T>
void launch_ewald_real(const VList& v, int na, int nb, Box sb, double rcut, double alpha, double* grad, double* parts,
                       hipStream_t s, int bpc, const PairTab* tab) {
    const double sfac = -2.0 * (alpha / sqrt(acos(-1.0)));
    const bool unit = sb.L[0] == 1.0 && sb.L[1] == 1.0 && sb.L[2] == 1.0;
    void (*k)(EWVL_PARAMS);
    if constexpr (sizeof(T) == 4)
        k = unit ? k_ewald_vl32<true> : k_ewald_vl32<false>;
    else
        k = unit ? k_ewald_vl64<true> : k_ewald_vl64<false>;
    static size_t dyn[2][2] = {{~(size_t)0, ~(size_t)0}, {~(size_t)0, ~(size_t)0}};  // [unit][capped]
    size_t& d = dyn[unit][bpc > 0];
    if (d == ~(size_t)0) d = sizeof(T) == 4 ? lds_for_blocks_per_cu((const void*)k, bpc) : 0;
    PairTab none{};
    const PairTab& t = tab ? *tab : none;
    const size_t dl = sizeof(T) == 4 ? d : 0;
    k<<<ewald_real_parts(na) * nb, 256, dl, s>>>(v.srt, v.srf, v.perm, v.cnt, v.nex, v.nl, na, nb, v.cap, sb.L[0],
                                                 sb.L[1], sb.L[2], sb.inv[0], sb.inv[1], sb.inv[2], rcut * rcut,
                                                 alpha, sfac, grad, parts, v.dmax, t.d, t.c, t.ku, t.kr);
}
template void launch_ewald_real<double>(const VList&, int, int, Box, double, double, double*, double*, hipStream_t,
                                        int, const PairTab*);
template void launch_ewald_real<float>(const VList&, int, int, Box, double, double, double*, double*, hipStream_t,
                                       int, const PairTab*);

// Smeared charges (src/ewald_smeared.f90:1-453): a pair of sites of different molecules interacts as
// two Gaussian charges, erfc(alpha r)/r - erfc(w r)/r with the A&S erfc in both terms, w by the
// sites' types (the first site of a molecule, M, and the two hydrogens, H): w0 for M-M, w1 for H-H,
// w2 for M-H.  The point-charge part is the real-space kernel's; this kernel, run after it on the
// same stream over the same list and block layout, adds the -erfc(w r)/r part in FP64 (both
// precision modes) to each site's gradient and to the block's energy and virial sums:
// du_s = -erfc(w r)/r, dur_s = (-sfac_w e^{-w^2 r^2} - du_s)/r^2, sfac_w = -2 w/sqrt(pi), the
// reference's expressions split by linearity.  A site's VLS lanes walk its list entries n = sl,
// sl + VLS, ... past its own-molecule ones; the lanes' sums are added in a fixed butterfly.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_ewald_smear(const double4* __restrict__ srt, const int* __restrict__ perm,
                                                     const int* __restrict__ cnt, const int* __restrict__ nex,
                                                     const int* __restrict__ nl, int na, int nb, double L0, double L1,
                                                     double L2, double i0, double i1, double i2, double rc2,
                                                     double w0, double w1, double w2, double* __restrict__ grad,
                                                     double* __restrict__ parts) {
    const int id = blockIdx.x;
    const int m8 = id >> 3;
    const int b = m8 % nb, sb = (id & 7) + 8 * (m8 / nb);
    const int k = (sb * blockDim.x + threadIdx.x) / VLS, sl = threadIdx.x % VLS;
    const double rfac = -2.0 / sqrt(acos(-1.0));
    double tg[3] = {0.0, 0.0, 0.0}, tv[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (k < na) {
        const double4* X = srt + (size_t)b * (na + 1);
        const double4 xi = X[k];
        const bool mi = perm[k] % 3 == 0;
        const int ntot = cnt[k];
        // the lane's first entry at or past the own-molecule ones
        int n = nex[k];
        n += (sl - n % VLS + VLS) % VLS;
        for (; n < ntot; n += VLS) {
            const int j = nl[(((size_t)(n / VLG) * na + k) * VLS + sl) * 4 + (n / VLS) % 4] >> 4;
            const double4 xj = X[j];
            double ex = xi.x - xj.x, ey = xi.y - xj.y, ez = xi.z - xj.z;
            if (UNIT) {
                ex -= rint(ex);
                ey -= rint(ey);
                ez -= rint(ez);
            } else {
                ex -= L0 * rint(ex * i0);
                ey -= L1 * rint(ey * i1);
                ez -= L2 * rint(ez * i2);
            }
            const double r2 = ex * ex + ey * ey + ez * ez;
            if (r2 >= rc2) continue;
            const bool mj = perm[j] % 3 == 0;
            const double wt = mi == mj ? (mi ? w0 : w1) : w2;
            const double rinv = 1.0 / sqrt(r2);
            const double xs = wt * (r2 * rinv);
            const double es = exp(-xs * xs);
            const double ts = AS_P / (AS_P + xs);
            const double erfcs = es * ts * (AS_A1 + ts * (AS_A2 + ts * (AS_A3 + ts * (AS_A4 + ts * AS_A5))));
            const double du = -erfcs * rinv;
            const double dur = (-(wt * rfac) * es - du) * (rinv * rinv);
            const double qij = xi.w * xj.w;
            const double dvr = qij * dur;
            const double fx = dvr * ex, fy = dvr * ey, fz = dvr * ez;
            tg[0] += fx;
            tg[1] += fy;
            tg[2] += fz;
            tv[0] += qij * du;
            tv[1] += ex * fx;
            tv[2] += ex * fy;
            tv[3] += ex * fz;
            tv[4] += ey * fy;
            tv[5] += ey * fz;
            tv[6] += ez * fz;
        }
    }
    // every lane of a quad takes part (the quad shares k)
#pragma unroll
    for (int c = 0; c < 3; c++) tg[c] = quad_sum(tg[c]);
#pragma unroll
    for (int c = 0; c < 7; c++) tv[c] = quad_sum(tv[c]);
    const bool own = k < na && sl == 0;
    if (own) {
        const int a = perm[k];
        double* G = grad + (size_t)b * 3 * na;
        G[a] += tg[0];
        G[na + a] += tg[1];
        G[2 * na + a] += tg[2];
    }
    // each pair is walked from both ends: half of the sums
    double acc[7];
#pragma unroll
    for (int c = 0; c < 7; c++) acc[c] = own ? 0.5 * tv[c] : 0.0;
    double w[10] = {acc[0], acc[1], acc[2], acc[3], acc[2], acc[4], acc[5], acc[3], acc[5], acc[6]};
    __shared__ double blk[10];
    block_reduce_store<10>(w, blk);
    // the thread that wrote blk[t] adds it to the real-space kernel's sum of this block
    if (threadIdx.x < 10) parts[((size_t)b * ewald_real_parts(na) + sb) * 10 + threadIdx.x] += blk[threadIdx.x];
}

void launch_ewald_smear(const VList& v, int na, int nb, Box sb, double rcut, const double wid[3], double* grad,
                        double* parts, hipStream_t s) {
    const bool unit = sb.L[0] == 1.0 && sb.L[1] == 1.0 && sb.L[2] == 1.0;
    auto k = unit ? k_ewald_smear<true> : k_ewald_smear<false>;
    k<<<ewald_real_parts(na) * nb, 256, 0, s>>>(v.srt, v.perm, v.cnt, v.nex, v.nl, na, nb, sb.L[0], sb.L[1], sb.L[2],
                                                sb.inv[0], sb.inv[1], sb.inv[2], rcut * rcut, wid[0], wid[1], wid[2],
                                                grad, parts);
}

// ------------------------------------------------------------------------------------------------
// reciprocal-space Ewald (src/ewald_cubic.f90:235-533; src/ewald_noncubic.f90:314-457)
//
// k tile: ky, kz in [-16, 15] (index k + 16), kx in [0, kmax], kmax <= 15.
// k_sites_vl: per (bead, atom) the base phases e^{i lat_d s_d}.
// k_sk_mfma : S(kx,ky,kz) = sum_j Y_j(ky) [q_j X_j(kx) Z_j(kz)], a complex GEMM over atoms on FP64
//             MFMA (v_mfma_f64_16x16x4f64).  A block owns KXB values of kx and one atom chunk; its
//             four waves own the four 16x16 (ky,kz) quadrants.  NJ atoms at a time are staged in LDS
//             as Y and Z tables (shared by the KXB values of kx) and X per kx.
// k_sk_final: sums the chunks, applies the reference weights w = (f/k^2) exp(-b k^2) on the
//             reference k set (mask built on the host with the reference's own comparisons), adds
//             energy and virial, stores G = 2 w S.
// k_kforce  : per atom, U(ky) = sum_kz G w^kz with w = conj(exp(i lat_z z)) by Horner's rule and its
//             kz-weighted derivative, two ky rows interleaved; kx split over grid.z into partial
//             gradients summed (in fixed order) by the chain-rule kernel.
// Mixed mode: k_sk_h16 and k_kforce_h16 take the same two contractions as GEMMs on FP16 MFMA with
// split operands (below); k_sk_final is shared.
// ------------------------------------------------------------------------------------------------
#define KT 32
#define KO 16
#define KH 15
#define NJ 32
#define SK_CHUNK 256
#define KXB 4
#define KXF 4

// the k set on the device for tiles of TT entries per ky and kz row (index k + TT / 2) and kx up to HH
template <int TT, int HH>
struct KDevT {
    static constexpr int T = TT, O = TT / 2, H = HH;
    int kmax;
    int cubic;
    int kyb[HH + 1];   // per kx: max |ky| with a member (-1: none)
    int kzb[HH + 1];   // per kx: max |kz| with a member
    signed char kzr[HH + 1][TT];  // per (kx, ky + O): max |kz| with a member (-1: none)
    double lat0, lat1, lat2;
    double f0, b;
};
typedef KDevT<KT, KH> KDev;
// kmax 16 .. 31 (FP64 in both precision modes): 64 x 64 tiles, the structure factor formed by the
// 32 x 32 kernel over four windows of the tile
typedef KDevT<2 * KT, 2 * KH + 1> KDevBig;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// z^n, n >= 0, by binary powering
__device__ __forceinline__ double2 cpow(double2 z, int n) {
    double2 r = make_double2(1.0, 0.0);
    while (n > 0) {
        if (n & 1) r = cmul(r, z);
        z = cmul(z, z);
        n >>= 1;
    }
    return r;
}

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// 16x16x4 FP64 MFMA; D layout (lane l, item v; lk = l >> 4, column l & 15): row lk + 4 v.  A and B:
// lane l gives A[l % 16][l / 16], B[l / 16][l % 16].  (The FP32-accumulating 16x16 MFMAs below
// have row 4 lk + v.)
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef dbl4 acc;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};

// T: MFMA precision (FP64; the mixed mode uses k_sk_h16).  KD: the k set's tile; the block forms the
// 32 x 32 window of it with ky from oy and kz from oz (the whole tile, oy = oz = -16, for kmax <= 15)
template <typename T, typename KD>
__global__ __launch_bounds__(256) void k_sk_mfma(const double2* __restrict__ ph, const double* __restrict__ q, int na,
                                                 int nchunk, KD kd, int oy, int oz, double2* __restrict__ skp) {
    constexpr int TT = KD::T, TO = KD::O;
    typedef typename Mfma<T>::acc acc_t;
    const int kxg = blockIdx.x, ch = blockIdx.y, b = blockIdx.z;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int qa = w >> 1, qc = w & 1;   // ky half, kz half of this wave
    const int kmax = kd.kmax;
    __shared__ T Yr[NJ][KT], Yi[NJ][KT], Zr[NJ][KT], Zi[NJ][KT];
    __shared__ T Xr[KXB][NJ], Xi[KXB][NJ];
    acc_t accR[KXB], accI[KXB];
#pragma unroll
    for (int g = 0; g < KXB; g++) {
        accR[g] = (acc_t){0, 0, 0, 0};
        accI[g] = (acc_t){0, 0, 0, 0};
    }
    const double2* PX = ph + ((size_t)b * 3 + 0) * na;
    const double2* PY = ph + ((size_t)b * 3 + 1) * na;
    const double2* PZ = ph + ((size_t)b * 3 + 2) * na;
    const int a0 = ch * SK_CHUNK, a1 = min(na, a0 + SK_CHUNK);
    for (int base = a0; base < a1; base += NJ) {
        {
            // Y and Z rows: thread -> (atom j = t/8, entries 4*(t%8) .. +3)
            const int j = t >> 3, part = t & 7, a = base + j;
            const int ky0 = 4 * part + oy, kz0 = 4 * part + oz;
            double2 y1 = make_double2(1.0, 0.0), z1 = make_double2(1.0, 0.0);
            const bool in = a < a1;
            if (in) {
                y1 = PY[a];
                z1 = PZ[a];
            }
            double2 yk = ky0 >= 0 ? cpow(y1, ky0) : cpow(make_double2(y1.x, -y1.y), -ky0);
            double2 zk = kz0 >= 0 ? cpow(z1, kz0) : cpow(make_double2(z1.x, -z1.y), -kz0);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int ky = ky0 + m, kz = kz0 + m;
                const bool oky = in && ky >= -kmax && ky <= kmax, okz = in && kz >= -kmax && kz <= kmax;
                Yr[j][4 * part + m] = oky ? (T)yk.x : (T)0;
                Yi[j][4 * part + m] = oky ? (T)yk.y : (T)0;
                Zr[j][4 * part + m] = okz ? (T)zk.x : (T)0;
                Zi[j][4 * part + m] = okz ? (T)zk.y : (T)0;
                yk = cmul(yk, y1);
                zk = cmul(zk, z1);
            }
            if (t < KXB * NJ) {
                const int g = t / NJ, jj = t % NJ, aa = base + jj, kx = kxg * KXB + g;
                double2 xv = make_double2(0.0, 0.0);
                if (aa < a1 && kx <= kmax) {
                    const double2 x1 = PX[aa];
                    const double2 p = cpow(x1, kx);
                    xv = make_double2(q[aa] * p.x, q[aa] * p.y);
                }
                Xr[g][jj] = (T)xv.x;
                Xi[g][jj] = (T)xv.y;
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int st = 0; st < NJ / 4; st++) {
            const int j = 4 * st + lk;
            const T ar = Yr[j][16 * qa + li], ai = Yi[j][16 * qa + li];
            const T zr = Zr[j][16 * qc + li], zi = Zi[j][16 * qc + li];
#pragma unroll
            for (int g = 0; g < KXB; g++) {
                const T xr = Xr[g][j], xi = Xi[g][j];
                const T br = xr * zr - xi * zi, bi = xr * zi + xi * zr;
                accR[g] = Mfma<T>::op(ar, br, accR[g]);
                accR[g] = Mfma<T>::op(-ai, bi, accR[g]);
                accI[g] = Mfma<T>::op(ar, bi, accI[g]);
                accI[g] = Mfma<T>::op(ai, br, accI[g]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int g = 0; g < KXB; g++) {
        const int kx = kxg * KXB + g;
        if (kx > kmax) break;
        double2* o = skp + (((size_t)b * (kmax + 1) + kx) * nchunk + ch) * (TT * TT);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int iy = 16 * qa + Mfma<T>::row(lk, v) + oy + TO, iz = 16 * qc + li + oz + TO;
            o[iy * TT + iz] = make_double2((double)accR[g][v], (double)accI[g][v]);
        }
    }
}

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __fp16 hp2 __attribute__((ext_vector_type(2)));
#define HS 36          // dwords per LDS operand row: 32 entries + 4 (16 rows x 16 B reads hit distinct banks)
#define KFH 64         // atoms per block of the force pass
#define KPAD 33        // padded FP32 LDS row

// (a, b) -> hi = fp16 pair rounded toward zero, lo = fp16 pair of (a, b) - hi (rounded to nearest).
// The mixed-precision FMA takes the difference straight from the packed halves of hi and rounds it
// to fp16 into one half of lo: three instructions per pair.
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
    uint32_t r;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(r) : "v"(h), "v"(b));
    hi = h;
    lo = r;
}
// (re, -im) -> (im, re) in a dword
__device__ __forceinline__ uint32_t conj_swap1(uint32_t v) { return __builtin_amdgcn_alignbit(v, v, 16) ^ 0x8000u; }
__device__ __forceinline__ uint4 conj_swap(uint4 v) {
    return make_uint4(conj_swap1(v.x), conj_swap1(v.y), conj_swap1(v.z), conj_swap1(v.w));
}
__device__ __forceinline__ flt4 mfma_h16(uint4 a, uint4 b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
}

// one block of 256 threads per (kx, bead), four (iy, iz) entries per thread: small blocks, so that the
// kernel is not held back while the real-space pairs occupy the CUs beside it; P: the partial sums'
// type.  SPLIT
// (mixed mode): instead of G, the A operand of k_kforce_h16 -- (G_r, -G_i) s as split half2 rows,
// rows 32..63 weighted by kz, hi then lo ([2][64][32] dwords per (bead, kx)) -- and 1 / (s 2^10) in
// gks, where the power of two s brings the largest |component| of the tile to [2^10, 2^11)
template <typename P, bool SPLIT, typename KD = KDev>
__device__ __forceinline__ void sk_final_body(const P* __restrict__ skp, int nchunk, const KD& kd,
                                              const unsigned char* __restrict__ mask, double2* __restrict__ gk,
                                              double* __restrict__ gks, double* __restrict__ parts, const int kx,
                                              const int b) {
    // TT x TT entries of the tile, four per thread (blocks of TT * TT / 4 threads)
    constexpr int TT = KD::T, TO = KD::O, NT = TT * TT / 4;
    static_assert(!SPLIT || TT == KT, "the split force-pass tiles are 32 x 32");
    const int kmax = kd.kmax;
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    const P* src = skp + ((size_t)b * (kmax + 1) + kx) * nchunk * (TT * TT);
    // thread -> entries e = threadIdx.x + NT i (iy = e / TT, iz = e % TT); each entry sums its chunks in
    // chunk order, the loads of four entries and up to four chunks in flight together
    double2 S[4];
#pragma unroll
    for (int i = 0; i < 4; i++) S[i] = make_double2(0.0, 0.0);
    int ch = 0;
    for (; ch + 4 <= nchunk; ch += 4) {
        P v[4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int i = 0; i < 4; i++) v[c][i] = src[(size_t)(ch + c) * (TT * TT) + threadIdx.x + NT * i];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                S[i].x += (double)v[c][i].x;
                S[i].y += (double)v[c][i].y;
            }
    }
    for (; ch < nchunk; ch++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const P v = src[(size_t)ch * (TT * TT) + threadIdx.x + NT * i];
            S[i].x += (double)v.x;
            S[i].y += (double)v.y;
        }
    double2 G[4];
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int e = threadIdx.x + NT * i, iy = e / TT, iz = e % TT;
        G[i] = make_double2(0.0, 0.0);
        if (mask[(kx * TT + iy) * TT + iz]) {
            const int ky = iy - TO, kz = iz - TO;
            const double rkx = kd.lat0 * kx, rky = kd.lat1 * ky, rkz = kd.lat2 * kz;
            const double rk2 = rkx * rkx + rky * rky + rkz * rkz;
            const double f = kx >= 1 ? 2.0 * kd.f0 : kd.f0;
            const double w = (f / rk2) * exp(-kd.b * rk2);
            const double et = w * (S[i].x * S[i].x + S[i].y * S[i].y);
            const double term = 2.0 * (1.0 / rk2 + kd.b);
            // The cubic reference sums the (+-ky, +-kz) members of one loop iteration into one et and
            // multiplies by the positive rky, rkz (src/ewald_cubic.f90:362-371,415-430,488-503), so its
            // off-diagonal virial uses |ky|, |kz|; the non-cubic sum keeps the signs
            // (src/ewald_noncubic.f90:426-441).
            const double rk[3] = {rkx, kd.cubic ? fabs(rky) : rky, kd.cubic ? fabs(rkz) : rkz};
            acc[0] += et;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) acc[1 + c * 3 + r] += et * (term * rk[r] * rk[c] - (r == c ? 1.0 : 0.0));
            G[i] = make_double2(2.0 * w * S[i].x, 2.0 * w * S[i].y);
        }
        m = fmax(m, fmax(fabs(G[i].x), fabs(G[i].y)));
    }
    if constexpr (!SPLIT) {
#pragma unroll
        for (int i = 0; i < 4; i++) gk[((size_t)b * (kmax + 1) + kx) * (TT * TT) + threadIdx.x + NT * i] = G[i];
    } else {
        __shared__ double wmax[4];
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
        __syncthreads();
        m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        int ex = 0;
        frexp(m, &ex);
        const double sc = m > 0.0 ? ldexp(1.0, 11 - ex) : 1.0;
        uint32_t* o = (uint32_t*)gk + ((size_t)b * (kmax + 1) + kx) * (4 * KT * KT);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = threadIdx.x + 256 * i, iy = e >> 5, iz = e & 31;
            const float gr = (float)(sc * G[i].x), gi = (float)(sc * G[i].y), kz = (float)(iz - KO);
            split_pair(gr, -gi, o[iy * KT + iz], o[2 * KT * KT + iy * KT + iz]);
            split_pair(kz * gr, -(kz * gi), o[(KT + iy) * KT + iz], o[2 * KT * KT + (KT + iy) * KT + iz]);
        }
        if (threadIdx.x == 0) gks[(size_t)b * (kmax + 1) + kx] = 1.0 / (sc * 1024.0);
    }
    block_reduce_store<10>(acc, parts + ((size_t)b * (kmax + 1) + kx) * 10);
}
template <typename P, bool SPLIT>
__global__ __launch_bounds__(256) void k_sk_final(const P* __restrict__ skp, int nchunk, KDev kd,
                                                  const unsigned char* __restrict__ mask, double2* __restrict__ gk,
                                                  double* __restrict__ gks, double* __restrict__ parts) {
    sk_final_body<P, SPLIT>(skp, nchunk, kd, mask, gk, gks, parts, blockIdx.x, blockIdx.y);
}
// kmax 16 .. 31: the 64 x 64 tiles in FP64, 1024 threads per (kx, bead)
__global__ __launch_bounds__(1024) void k_sk_final_big(const double2* __restrict__ skp, int nchunk, KDevBig kd,
                                                       const unsigned char* __restrict__ mask,
                                                       double2* __restrict__ gk, double* __restrict__ parts) {
    sk_final_body<double2, false, KDevBig>(skp, nchunk, kd, mask, gk, nullptr, parts, blockIdx.x, blockIdx.y);
}

// Horner over kz of two ky rows: h = sum_m c_m w^m, d = dh/dw, with c_m = G(kz = m - KZ).  After the
// leading coefficient there are 2 KZ steps, taken two at a time with the four coefficients loaded
// first, so the loads overlap the FMA chains.
__device__ __forceinline__ void horner_step(double2& h, double2& d, double2 w, double2 g) {
    d = make_double2(fma(d.x, w.x, fma(-d.y, w.y, h.x)), fma(d.x, w.y, fma(d.y, w.x, h.y)));
    h = make_double2(fma(h.x, w.x, fma(-h.y, w.y, g.x)), fma(h.x, w.y, fma(h.y, w.x, g.y)));
}
template <int TO = KO>
__device__ __forceinline__ void horner2(const double2* __restrict__ ra, const double2* __restrict__ rb, int KZ,
                                        double2 w, double2& ha, double2& da, double2& hb, double2& db) {
    ha = ra[KZ + TO];
    hb = rb[KZ + TO];
    da = make_double2(0.0, 0.0);
    db = make_double2(0.0, 0.0);
    for (int kz = KZ - 1; kz >= -KZ; kz -= 2) {
        const double2 ga0 = ra[kz + TO], gb0 = rb[kz + TO];
        const double2 ga1 = ra[kz - 1 + TO], gb1 = rb[kz - 1 + TO];
        horner_step(ha, da, w, ga0);
        horner_step(hb, db, w, gb0);
        horner_step(ha, da, w, ga1);
        horner_step(hb, db, w, gb1);
    }
}

// one thread per (bead, atom, kx group); writes the group's partial gradient (unit-box / physical)
template <typename KD>
__global__ __launch_bounds__(256) void k_kforce(const double2* __restrict__ ph, const double* __restrict__ q, int na,
                                                KD kd, const double2* __restrict__ gk, double* __restrict__ kpart) {
    constexpr int TT = KD::T, TO = KD::O;
    const int b = blockIdx.y, g = blockIdx.z, nb = gridDim.y;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int kmax = kd.kmax;
    const double2 X1 = ph[((size_t)b * 3 + 0) * na + a], Y1 = ph[((size_t)b * 3 + 1) * na + a];
    const double2 Z1 = ph[((size_t)b * 3 + 2) * na + a];
    const double2 w = make_double2(Z1.x, -Z1.y), Yc = make_double2(Y1.x, -Y1.y);
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const int kx0 = g * KXF, kx1 = min(kmax, kx0 + KXF - 1);
    double2 xk = cpow(X1, kx0);
    xk = make_double2(q[a] * xk.x, q[a] * xk.y);
    const double2* G0 = gk + (size_t)b * (kmax + 1) * TT * TT;
    for (int kx = kx0; kx <= kx1; kx++) {
        if (kx > kx0) xk = cmul(xk, X1);
        const int KY = kd.kyb[kx];
        if (KY < 0) continue;
        double2 yc = cpow(Yc, KY);   // Y(-KY)
        double tx = 0.0, ty = 0.0, tz = 0.0;
        const double2* Gx = G0 + (size_t)kx * TT * TT;
        int ky = -KY;
        for (; ky + 1 <= KY; ky += 2) {
            // the pair's kz range: the larger of the two rows' bounds (G is zero beyond)
            const int KZ = max((int)kd.kzr[kx][ky + TO], (int)kd.kzr[kx][ky + 1 + TO]);
            const double2 yb = cmul(yc, Y1);
            if (KZ >= 0) {
                const double2 zK = cpow(Z1, KZ);
                double2 ha, da, hb, db;
                horner2<TO>(Gx + (ky + TO) * TT, Gx + (ky + 1 + TO) * TT, KZ, w, ha, da, hb, db);
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const double2 h = r ? hb : ha, d = r ? db : da, yy = r ? yb : yc;
                    const double2 U = cmul(zK, h);
                    const double2 wd = cmul(w, d);
                    const double2 Uz = cmul(zK, make_double2(wd.x - KZ * h.x, wd.y - KZ * h.y));
                    const double2 xy = cmul(xk, yy);
                    const double ipu = xy.x * U.y - xy.y * U.x;   // Im(conj(xy) U)
                    const double ipz = xy.x * Uz.y - xy.y * Uz.x;
                    tx += ipu;
                    ty += (ky + r) * ipu;
                    tz += ipz;
                }
            }
            yc = cmul(yb, Y1);
        }
        const int KZ = ky <= KY ? (int)kd.kzr[kx][ky + TO] : -1;
        if (KZ >= 0) {
            const double2 zK = cpow(Z1, KZ);
            double2 ha, da, hb, db;
            horner2<TO>(Gx + (ky + TO) * TT, Gx + (ky + TO) * TT, KZ, w, ha, da, hb, db);
            const double2 U = cmul(zK, ha);
            const double2 wd = cmul(w, da);
            const double2 Uz = cmul(zK, make_double2(wd.x - KZ * ha.x, wd.y - KZ * ha.y));
            const double2 xy = cmul(xk, yc);
            const double ipu = xy.x * U.y - xy.y * U.x;
            const double ipz = xy.x * Uz.y - xy.y * Uz.x;
            tx += ipu;
            ty += ky * ipu;
            tz += ipz;
        }
        gx += kd.lat0 * kx * tx;
        gy += kd.lat1 * ty;
        gz += kd.lat2 * tz;
    }
    double* o = kpart + ((size_t)g * nb + b) * 3 * na;
    o[a] = gx;
    o[na + a] = gy;
    o[2 * na + a] = gz;
}

// ---- mixed mode: the k-space GEMMs on FP16 MFMA with split operands -----------------------------
// Each FP32 operand v is carried as two FP16 numbers, hi = fp16(v) and lo = fp16(v - hi), so that
// hi + lo holds about 21 significant bits, and a product is taken as hi*hi + hi*lo + lo*hi (the
// split-FP16 product of H. Ootomo and R. Yokota, "Recovering single precision accuracy from Tensor
// Cores while surpassing the FP32 theoretical peak performance", IJHPCA 36 (2022)).  The three
// products share one FP32 accumulator (the paper keeps the small ones apart, against the truncating
// accumulation of its hardware); here the force error against the FP64 path measured the same with
// one accumulator as with two (3.8e-7 relative RMS at 1728 H2O).  The tables are scaled by powers of two that keep
// the lo parts normal FP16 numbers; hi is rounded toward zero (v_cvt_pkrtz), so v - hi is exact in
// FP32.  On v_mfma_f32_16x16x32_f16 the three products cost about a fifth of one FP32 MFMA product.
//
// A complex contraction sum_j a_j b_j is one real GEMM over an interleaved K: slot 2j holds the real
// and slot 2j + 1 the imaginary part, so one dword of the operand tables is the half2 (re, im) of one
// entry.  With B = (b_r, b_i) per entry, Re = A_r.B with A_r = (a_r, -a_i) and Im = A_i.B with
// A_i = (a_i, a_r), and A_i follows from A_r by swapping the halves of each dword and negating the
// new low half.  The f16 MFMA pairs element e of lane group g of A with element e of lane group g of
// B, so any assignment of K slots to lane elements that A and B share gives the same sum.
// z^e for |e| <= 16 with e a multiple of m (m = 4 or 8: the bits below m are zero), branch-free
__device__ __forceinline__ double2 cpow_sel(double2 z, int e, int m) {
    const double2 one = make_double2(1.0, 0.0);
    const int n = abs(e);
    if (e < 0) z = make_double2(z.x, -z.y);
    double2 p = z;
    for (int s = 1; s < m; s <<= 1) p = cmul(p, p);  // z^m
    double2 r = one;
    for (int s = m; s <= 16; s <<= 1) {
        r = cmul(r, (n & s) ? p : one);
        p = cmul(p, p);
    }
    return r;
}

// S(kx, ky, kz) = sum_j Y_j(ky) W_j(kz), W_j = q_j X_j^kx Z_j^kz.  One block of 512 threads per
// (atom chunk, bead) owns every kx: wave w owns the kz half w & 1, kx = 4 (w >> 1) .. + 3 and both ky
// halves.  Per SKB atoms the block forms, from FP64 powers of the base phases, the A operands
// (Y_r, -Y_i) x 2^10 and (Y_i, Y_r) x 2^10 [ky][atom] as split half2, Z^kz [kz][atom] and
// q X^kx x 2^10 [kx][atom] in FP32, in one of two LDS buffers (one barrier per SKB atoms: the next
// tables are formed while the MFMAs read the current ones).  Each lane forms its B operand, W for
// its kz and its four atoms of a 16-atom step, in registers per kx and splits it: no W goes through
// LDS.  Per 16 atoms a wave issues 4 kx x 2 ky halves x (Re, Im) x 3 MFMAs.  Partial sums per
// chunk in FP32.
#define SKB 32
// scheduling hint for the straight-line block before it: N groups of one MFMA followed by V other
// vector instructions, so that the vector work issues while the matrix unit runs
template <int N, int V>
__device__ __forceinline__ void mfma_interleave() {
#pragma unroll
    for (int i = 0; i < N; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, V, 0);
    }
}
__device__ __forceinline__ void sk_h16_body(const double2* __restrict__ ph, const double* __restrict__ q, int na,
                                            int chunk, int nchunk, const KDev& kd, float2* __restrict__ skp,
                                            const int ch, const int b) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int qc = w & 1, kg = w >> 1;
    const int kmax = kd.kmax;
    __shared__ __align__(16) uint32_t Ab[2][4][KT][HS];  // [buffer][(r, i) x (hi, lo)][ky][atom]
    __shared__ __align__(16) float2 Zb[2][KT][SKB + 2];   // [buffer][kz][atom]
    __shared__ __align__(16) float2 Xb[2][KH + 1][SKB];   // [buffer][kx][atom]
    flt4 sr[4][2], si[4][2];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int h = 0; h < 2; h++) sr[g][h] = si[g][h] = (flt4){0, 0, 0, 0};
    const double2* PX = ph + ((size_t)b * 3 + 0) * na;
    const double2* PY = ph + ((size_t)b * 3 + 1) * na;
    const double2* PZ = ph + ((size_t)b * 3 + 2) * na;
    const int a0 = ch * chunk, a1 = min(na, a0 + chunk);
    constexpr double SC = 1024.0;
    // this thread's base phases (Y or Z, and X) and charge of an atom, loaded a stage ahead of use;
    // loads past the chunk are clamped into it and their values unused
    const int j = t & 31;
    const double2* PV = t < 256 ? PY : PZ;
    auto load = [&](int base, double2& v, double2& x, double& qv) {
        const int aa = min(base + j, a1 - 1);
        v = PV[aa];
        x = PX[aa];
        qv = q[aa];
    };
    // tables of the SKB atoms from base into buffer u
    auto form = [&](int base, int u, double2 v1, double2 x1, double qa) {
        const bool in = base + j < a1;
        const double2 one = make_double2(1.0, 0.0);
        if (t < 256) {
            // Y: thread -> (ky = 4 (t / 32) - 16 .. + 3, atom t % 32)
            const int k0 = 4 * (t >> 5) - KO;
            const double2 y1 = in ? v1 : one;
            double2 yk = cpow_sel(y1, k0, 4);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int k = k0 + m;
                const bool ok = in && k >= -kmax && k <= kmax;
                uint32_t h, lo;
                split_pair(ok ? (float)(SC * yk.x) : 0.0f, ok ? -(float)(SC * yk.y) : 0.0f, h, lo);
                Ab[u][0][k + KO][j] = h;
                Ab[u][1][k + KO][j] = lo;
                Ab[u][2][k + KO][j] = conj_swap1(h);
                Ab[u][3][k + KO][j] = conj_swap1(lo);
                yk = cmul(yk, y1);
            }
        } else {
            // Z: thread -> (kz = 4 ((t - 256) / 32) - 16 .. + 3, atom t % 32)
            const int k0 = 4 * ((t - 256) >> 5) - KO;
            const double2 z1 = in ? v1 : one;
            double2 zk = cpow_sel(z1, k0, 4);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int k = k0 + m;
                const bool ok = in && k >= -kmax && k <= kmax;
                Zb[u][k + KO][j] = ok ? make_float2((float)zk.x, (float)zk.y) : make_float2(0.0f, 0.0f);
                zk = cmul(zk, z1);
            }
        }
        // q X^kx x 2^10: thread -> (kx = t / 32, atom t % 32)
        const int kx = t >> 5;
        double2 v = make_double2(0.0, 0.0);
        if (in && kx <= kmax) {
            double2 p = x1, r = one;
#pragma unroll
            for (int s = 1; s <= 8; s <<= 1) {
                if (kx & s) r = cmul(r, p);
                if (s < 8) p = cmul(p, p);
            }
            const double qs = SC * qa;
            v = make_double2(qs * r.x, qs * r.y);
        }
        Xb[u][kx][j] = make_float2((float)v.x, (float)v.y);
    };
    const int nst = (a1 - a0 + SKB - 1) / SKB;
    double2 pv, px;
    double pq;
    load(a0, pv, px, pq);
    form(a0, 0, pv, px, pq);
    load(a0 + SKB, pv, px, pq);
    __syncthreads();
    for (int st = 0; st < nst; st++) {
        const int u = st & 1;
        if (st + 1 < nst) {
            const double2 cv = pv, cx = px;
            const double cq = pq;
            load(a0 + (st + 2) * SKB, pv, px, pq);
            form(a0 + (st + 1) * SKB, u ^ 1, cv, cx, cq);
        }
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int c = 16 * ks + 4 * lk;  // this lane's four atoms of the 16-atom step
            uint4 arh[2], arl[2], aih[2], ail[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                arh[h] = *(const uint4*)&Ab[u][0][16 * h + li][c];
                arl[h] = *(const uint4*)&Ab[u][1][16 * h + li][c];
                aih[h] = *(const uint4*)&Ab[u][2][16 * h + li][c];
                ail[h] = *(const uint4*)&Ab[u][3][16 * h + li][c];
            }
            const float4 z01 = *(const float4*)&Zb[u][16 * qc + li][c];
            const float4 z23 = *(const float4*)&Zb[u][16 * qc + li][c + 2];
            const float2 z[4] = {make_float2(z01.x, z01.y), make_float2(z01.z, z01.w), make_float2(z23.x, z23.y),
                                 make_float2(z23.z, z23.w)};
            float4 xv[4][2];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                xv[g][0] = *(const float4*)&Xb[u][4 * kg + g][c];
                xv[g][1] = *(const float4*)&Xb[u][4 * kg + g][c + 2];
            }
            // W of kx 4 kg + g, split: formed one kx ahead of its MFMAs (X is zero past kmax)
            auto formw = [&](int g, uint4& bh, uint4& bl) {
                const float2 x[4] = {make_float2(xv[g][0].x, xv[g][0].y), make_float2(xv[g][0].z, xv[g][0].w),
                                     make_float2(xv[g][1].x, xv[g][1].y), make_float2(xv[g][1].z, xv[g][1].w)};
                uint32_t hi[4], lo[4];
#pragma unroll
                for (int i = 0; i < 4; i++)
                    split_pair(x[i].x * z[i].x - x[i].y * z[i].y, x[i].x * z[i].y + x[i].y * z[i].x, hi[i], lo[i]);
                bh = make_uint4(hi[0], hi[1], hi[2], hi[3]);
                bl = make_uint4(lo[0], lo[1], lo[2], lo[3]);
            };
            uint4 bh[2], bl[2];
            formw(0, bh[0], bl[0]);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                if (g + 1 < 4) formw(g + 1, bh[(g + 1) & 1], bl[(g + 1) & 1]);
                const uint4 h0 = bh[g & 1], l0 = bl[g & 1];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    sr[g][h] = mfma_h16(arh[h], h0, sr[g][h]);
                    si[g][h] = mfma_h16(aih[h], h0, si[g][h]);
                    sr[g][h] = mfma_h16(arh[h], l0, sr[g][h]);
                    si[g][h] = mfma_h16(aih[h], l0, si[g][h]);
                    sr[g][h] = mfma_h16(arl[h], h0, sr[g][h]);
                    si[g][h] = mfma_h16(ail[h], h0, si[g][h]);
                }
            }
        }
        __syncthreads();
    }
    const float us = (float)(1.0 / (SC * SC));
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int kx = 4 * kg + g;
        if (kx > kmax) break;
        float2* o = skp + (((size_t)b * (kmax + 1) + kx) * nchunk + ch) * (KT * KT);
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int iy = 16 * h + 4 * lk + v, iz = 16 * qc + li;
                o[iy * KT + iz] = make_float2(sr[g][h][v] * us, si[g][h][v] * us);
            }
    }
}
__global__ __launch_bounds__(512) void k_sk_h16(const double2* __restrict__ ph, const double* __restrict__ q, int na,
                                                int chunk, int nchunk, KDev kd, float2* __restrict__ skp) {
    sk_h16_body(ph, q, na, chunk, nchunk, kd, skp, blockIdx.x, blockIdx.y);
}

// The force pass as a GEMM: for each kx, U = [G; kz G] conj(Z), A = the (ky, kz) tile of G = 2 w S
// (rows 32..63 weighted by kz; split and scaled by k_sk_final) and B = conj(Z_j(kz)) for 64 atoms, so
// U(ky, j) = sum_kz G(ky, kz) e^{-i kz z_j} and its kz-weighted twin; per atom the sums over ky of
// Im(conj(q X^kx Y^ky) U) (and ky-weighted, and with the twin) give the three gradient components,
// the quantities of the Horner pass.  One block per (64 atoms, bead); wave w owns rows
// 16 w .. 16 w + 15 of [G; kz G], which it alone reads: its A operand comes from global memory
// straight into registers, two kx ahead, so the kx loop has no barrier.  B (x 2^10), Y^ky and
// q X^kx for every kx are formed once in LDS, with the tile scales, and read from there.  The MFMAs
// of kx + 1 are issued before the per-atom sums of kx (two accumulator sets), so a wave keeps the
// matrix and vector units busy together; at most 256 registers (two waves per SIMD).  Each lane
// sums its part of the three components over kx in FP32 (the tile scales are powers of two); lanes
// and waves are combined once at the end, in fixed order, the waves in FP64.  Blocks are dealt to the XCDs by bead.
__device__ __forceinline__ void kforce_h16_body(const double2* __restrict__ ph, const double* __restrict__ q, int na,
                                                int nbk, const KDev& kd, const uint4* __restrict__ ga,
                                                const double* __restrict__ gks, double* __restrict__ kpart,
                                                const int bid) {
    // blocks are dealt to the eight XCDs in turn (block id mod 8): XCD x takes beads x, x + 8, ... one
    // after the other, so the blocks resident on an XCD read the G tiles of one bead from its L2
    const int nab = (na + KFH - 1) / KFH, m8 = bid >> 3;
    const int b = (bid & 7) + 8 * (m8 / nab), a0 = (m8 % nab) * KFH;
    if (b >= nbk) return;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int kmax = kd.kmax;
    __shared__ uint32_t Zh[KFH][HS], Zl[KFH][HS];   // conj(Z_j(kz)) x 2^10 [atom][kz]
    __shared__ float2 Ys[KT][KFH + 2];              // Y_j(ky) [ky][atom]
    __shared__ float2 Xs[KH + 1][KFH];              // q X_j^kx [kx][atom]
    __shared__ double red[4][KFH][2];               // per wave: its two sums per atom
    __shared__ double gs[KH + 1];                   // 1 / (tile scale 2^10) per kx
    constexpr double SC = 1024.0;
    const double2* PX = ph + ((size_t)b * 3 + 0) * na;
    const double2* PY = ph + ((size_t)b * 3 + 1) * na;
    const double2* PZ = ph + ((size_t)b * 3 + 2) * na;
    {
        // thread -> (atom t / 4, entries 8 (t % 4) .. + 7)
        const int j = t >> 2, part = t & 3, a = a0 + j;
        const int k0 = 8 * part - KO;
        const bool in = a < na;
        const double2 y1 = in ? PY[a] : make_double2(1.0, 0.0), z1 = in ? PZ[a] : make_double2(1.0, 0.0);
        double2 yk = cpow_sel(y1, k0, 8), zk = cpow_sel(z1, k0, 8);
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int k = k0 + m;
            const bool ok = in && k >= -kmax && k <= kmax;
            Ys[k + KO][j] = ok ? make_float2((float)yk.x, (float)yk.y) : make_float2(0.0f, 0.0f);
            split_pair(ok ? (float)(SC * zk.x) : 0.0f, ok ? -(float)(SC * zk.y) : 0.0f, Zh[j][k + KO], Zl[j][k + KO]);
            yk = cmul(yk, y1);
            zk = cmul(zk, z1);
        }
        // q X^kx: thread -> (atom t / 4, kx = 4 (t % 4) .. + 3)
        const int kx0 = 4 * part;
        const double2 x1 = in ? PX[a] : make_double2(0.0, 0.0);
        double2 xk = cpow_sel(x1, kx0, 4);
        const double qq = in ? q[a] : 0.0;
        xk = make_double2(qq * xk.x, qq * xk.y);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            Xs[kx0 + m][j] = make_float2((float)xk.x, (float)xk.y);
            xk = cmul(xk, x1);
        }
    }
    if (t <= kmax) gs[t] = gks[(size_t)b * (kmax + 1) + t];
    __syncthreads();
    const int mb = w & 1, wz = w >> 1;  // row block of G, kz-weighted twin
    const int r0 = KT * wz + 16 * mb;
    // A operand of lane (li, lk) for kx: rows r0 + li, dwords 16 ks + 4 lk .. + 3, hi and lo
    const uint4* GA = ga + (size_t)b * (kmax + 1) * (KT * KT) + (r0 + li) * 8 + lk;
    float sa[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // a kx without members has a zero tile: its sums add zero
    auto mma = [&](uint4 ah0, uint4 al0, uint4 ah1, uint4 al1, flt4 (&ur)[4], flt4 (&ui)[4]) {
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const uint4 ah = ks ? ah1 : ah0, al = ks ? al1 : al0;
            const uint4 a2h = conj_swap(ah), a2l = conj_swap(al);
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const uint4 bh = *(const uint4*)&Zh[16 * n + li][16 * ks + 4 * lk];
                const uint4 bl = *(const uint4*)&Zl[16 * n + li][16 * ks + 4 * lk];
                if (ks == 0) {
                    ur[n] = mfma_h16(ah, bh, (flt4){0, 0, 0, 0});
                    ui[n] = mfma_h16(a2h, bh, (flt4){0, 0, 0, 0});
                } else {
                    ur[n] = mfma_h16(ah, bh, ur[n]);
                    ui[n] = mfma_h16(a2h, bh, ui[n]);
                }
                ur[n] = mfma_h16(ah, bl, ur[n]);
                ur[n] = mfma_h16(al, bh, ur[n]);
                ui[n] = mfma_h16(a2h, bl, ui[n]);
                ui[n] = mfma_h16(a2l, bh, ui[n]);
            }
        }
    };
    auto sums = [&](int kx, const flt4 (&ur)[4], const flt4 (&ui)[4]) {
        const float us = (float)gs[kx];
        const float wa = wz ? us : us * (float)kx;  // x: kx-weighted plain sum; z: the twin's plain sum
        // rows of this lane: iy = 16 mb + 4 lk + v (ky = iy - 16); column: atom 16 n + li
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int j = 16 * n + li;
            const float2 x = Xs[kx][j];
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int iy = 16 * mb + 4 * lk + v;
                const float2 y = Ys[iy][j];
                const float pr = x.x * y.x - x.y * y.y, pi = x.x * y.y + x.y * y.x;  // q X^kx Y^ky
                const float ipu = pr * ui[n][v] - pi * ur[n][v];                      // Im(conj(p) U)
                s0 += ipu;
                s1 += (float)(iy - KO) * ipu;
            }
            sa[n] += wa * s0;
            sb[n] += us * s1;
        }
    };
    // A rows: even kx in set 0, odd kx in set 1, each loaded two kx ahead (loads past kmax re-read
    // kmax's tile, unused)
    auto load = [&](int kx, uint4 (&A)[4]) {
        const uint4* g = GA + (size_t)min(kx, kmax) * (KT * KT);
        A[0] = g[0];
        A[1] = g[512];
        A[2] = g[4];
        A[3] = g[516];
    };
    uint4 A0[4], A1[4];
    flt4 urA[4], uiA[4], urB[4], uiB[4];
    load(0, A0);
    load(1, A1);
    mma(A0[0], A0[1], A0[2], A0[3], urA, uiA);
    load(2, A0);
    for (int kx = 0; kx <= kmax; kx += 2) {
        if (kx + 1 <= kmax) mma(A1[0], A1[1], A1[2], A1[3], urB, uiB);
        load(kx + 3, A1);
        sums(kx, urA, uiA);
        if (kx + 1 > kmax) break;
        if (kx + 2 <= kmax) mma(A0[0], A0[1], A0[2], A0[3], urA, uiA);
        load(kx + 4, A0);
        sums(kx + 1, urB, uiB);
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
        sa[n] += __shfl_xor(sa[n], 16, 64);
        sb[n] += __shfl_xor(sb[n], 16, 64);
        sa[n] += __shfl_xor(sa[n], 32, 64);
        sb[n] += __shfl_xor(sb[n], 32, 64);
        if (lk == 0) {
            red[w][16 * n + li][0] = sa[n];
            red[w][16 * n + li][1] = sb[n];
        }
    }
    __syncthreads();
    if (t < KFH && a0 + t < na) {
        // waves 0, 1: (kx-weighted, ky-weighted) sums of the two ky halves; waves 2, 3: the twin
        const double tx = red[0][t][0] + red[1][t][0];
        const double ty = red[0][t][1] + red[1][t][1];
        const double tz = red[2][t][0] + red[3][t][0];
        double* o = kpart + (size_t)b * 3 * na;
        o[a0 + t] = kd.lat0 * tx;
        o[na + a0 + t] = kd.lat1 * ty;
        o[2 * na + a0 + t] = kd.lat2 * tz;
    }
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_kforce_h16(
    const double2* __restrict__ ph, const double* __restrict__ q, int na, int nbk, KDev kd,
    const uint4* __restrict__ ga, const double* __restrict__ gks, double* __restrict__ kpart) {
    kforce_h16_body(ph, q, na, nbk, kd, ga, gks, kpart, blockIdx.x);
}

int kspace_nchunk(int na) { return (na + SK_CHUNK - 1) / SK_CHUNK; }
// mixed structure-factor pass: about one block per CU over all beads (the FP32 partials of at most
// twice the FP64 pass's chunks fit its buffer); chunks of a multiple of SKB atoms
static int kspace_nchunk_h16(int na, int nb, int* chunk) {
    int n = std::min(2 * kspace_nchunk(na), std::max(1, (256 + nb - 1) / nb));
    const int c = ((na + n - 1) / n + SKB - 1) / SKB * SKB;
    *chunk = c;
    return (na + c - 1) / c;
}
int kspace_ngroups(int kmax) { return (kmax + KXF) / KXF; }

static KDev make_kdev(const KSet& ks) {
    KDev kd;
    kd.kmax = ks.kmax;
    for (int kx = 0; kx <= KH; kx++) {
        kd.kyb[kx] = ks.ky_max[kx];
        kd.kzb[kx] = ks.kz_bound[kx];
        for (int iy = 0; iy < KT; iy++) kd.kzr[kx][iy] = ks.kz_row[kx][iy];
    }
    kd.lat0 = ks.lat[0];
    kd.lat1 = ks.lat[1];
    kd.lat2 = ks.lat[2];
    kd.f0 = ks.f0;
    kd.b = ks.b;
    kd.cubic = ks.cubic;
    return kd;
}

void launch_sk(const double* ph, const double* q, int na, int nb, const KSet& ks, double* sk_part, int fp32,
               hipStream_t s) {
    const KDev kd = make_kdev(ks);
    if (fp32) {
        int chunk = 0;
        const int nchunk = kspace_nchunk_h16(na, nb, &chunk);
        k_sk_h16<<<dim3(nchunk, nb), 512, 0, s>>>((const double2*)ph, q, na, chunk, nchunk, kd, (float2*)sk_part);
        return;
    }
    const int nchunk = kspace_nchunk(na);
    k_sk_mfma<double, KDev><<<dim3((ks.kmax + KXB) / KXB, nchunk, nb), 256, 0, s>>>(
        (const double2*)ph, q, na, nchunk, kd, -KO, -KO, (double2*)sk_part);
}
void launch_sk_final(int na, int nb, const KSet& ks, const unsigned char* dmask, const double* sk_part, double* gk,
                     double* gks, double* parts_k, int fp32, hipStream_t s) {
    const KDev kd = make_kdev(ks);
    const dim3 grid(ks.kmax + 1, nb);
    int chunk = 0;
    if (fp32)
        k_sk_final<float2, true><<<grid, 256, 0, s>>>((const float2*)sk_part, kspace_nchunk_h16(na, nb, &chunk), kd,
                                                       dmask, (double2*)gk, gks, parts_k);
    else
        k_sk_final<double2, false><<<grid, 256, 0, s>>>((const double2*)sk_part, kspace_nchunk(na), kd, dmask,
                                                         (double2*)gk, gks, parts_k);
}
// the force pass; returns the number of partial-gradient slots written (kx groups, or 1)
int launch_kforce(const double* ph, const double* q, int na, int nb, const KSet& ks, const double* gk,
                  const double* gks, double* kpart, int fp32, hipStream_t s) {
    const KDev kd = make_kdev(ks);
    if (fp32) {
        k_kforce_h16<<<(na + KFH - 1) / KFH * ((nb + 7) / 8 * 8), 256, 0, s>>>((const double2*)ph, q, na, nb, kd,
                                                                              (const uint4*)gk, gks, kpart);
        return 1;
    }
    k_kforce<KDev><<<dim3((na + 255) / 256, nb, kspace_ngroups(ks.kmax)), 256, 0, s>>>(
        (const double2*)ph, q, na, kd, (const double2*)gk, kpart);
    return kspace_ngroups(ks.kmax);
}

// kmax 16 .. 31 (both precision modes): the FP64 kernels on 64 x 64 tiles -- the structure factor
// over the tile's four 32 x 32 windows, the weights and sums, the force pass; returns the number of
// partial-gradient slots written.  skp: (kmax + 1) x chunks x 64 x 64 per bead, gk: (kmax + 1) x 64 x 64
// per bead (double2), parts_k: (kmax + 1) x 10 per bead.
int launch_kspace_big(const double* ph, const double* q, int na, int nb, const KSetBig& ks, const unsigned char* dmask,
                      double* skp, double* gk, double* parts_k, double* kpart, hipStream_t s) {
    KDevBig kd;
    kd.kmax = ks.kmax;
    for (int kx = 0; kx <= KDevBig::H; kx++) {
        kd.kyb[kx] = ks.ky_max[kx];
        kd.kzb[kx] = ks.kz_bound[kx];
        for (int iy = 0; iy < KDevBig::T; iy++) kd.kzr[kx][iy] = ks.kz_row[kx][iy];
    }
    kd.lat0 = ks.lat[0];
    kd.lat1 = ks.lat[1];
    kd.lat2 = ks.lat[2];
    kd.f0 = ks.f0;
    kd.b = ks.b;
    kd.cubic = ks.cubic;
    const int nchunk = kspace_nchunk(na);
    for (int oy = -KDevBig::O; oy < KDevBig::O; oy += KT)
        for (int oz = -KDevBig::O; oz < KDevBig::O; oz += KT)
            k_sk_mfma<double, KDevBig><<<dim3((ks.kmax + KXB) / KXB, nchunk, nb), 256, 0, s>>>(
                (const double2*)ph, q, na, nchunk, kd, oy, oz, (double2*)skp);
    k_sk_final_big<<<dim3(ks.kmax + 1, nb), 1024, 0, s>>>((const double2*)skp, nchunk, kd, dmask, (double2*)gk,
                                                          parts_k);
    k_kforce<KDevBig><<<dim3((na + 255) / 256, nb, kspace_ngroups(ks.kmax)), 256, 0, s>>>(
        (const double2*)ph, q, na, kd, (const double2*)gk, kpart);
    return kspace_ngroups(ks.kmax);
}

// ------------------------------------------------------------------------------------------------
// Replica batches: the four dominant force kernels of many contexts (the FP32 pair walk on the unit
// box, k_sk_h16, k_sk_final, k_kforce_h16), each as one launch over every replica's blocks.  A
// replica's blocks form a range starting at a multiple of 8 (so the XCD dealing of the pair and
// force-pass kernels, block id mod 8, is the solo launch's), its arguments are one record in HBM,
// and every block finds its replica by a binary search of the range starts.  The kernels' bodies
// are the solo kernels', with the same arguments, so each replica's results are the solo results
// bit for bit.
// ------------------------------------------------------------------------------------------------
struct EwVlArgs {
    const double4* srt;
    const uint4* srf;
    const int *perm, *cnt, *nex, *nl;
    int na, nb, cap, nblk;
    double L0, L1, L2, i0, i1, i2, rc2, alpha, sfac;
    double *grad, *parts;
    unsigned long long* dmax;
    const float4* tab;
    float tsc;
    int tku, tkr;
};
struct SkH16Args {
    const double2* ph;
    const double* q;
    float2* skp;
    int na, chunk, nchunk, nblk;
    KDev kd;
};
struct SkFinArgs {
    const float2* skp;
    const unsigned char* mask;
    double2* gk;
    double *gks, *parts;
    int nchunk, nblk;
    KDev kd;
};
struct KfH16Args {
    const double2* ph;
    const double* q;
    const uint4* ga;
    const double* gks;
    double* kpart;
    int na, nbk, nblk;
    KDev kd;
};

// the replica of block bid: the last r with off[r] <= bid (off[0] = 0, n ranges)
__device__ __forceinline__ int batch_find(const int* __restrict__ off, int n, int bid) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= bid)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256, 8) void k_ewald_vl32_batch(const EwVlArgs* __restrict__ A,
                                                             const int* __restrict__ off, int n) {
    const int r = batch_find(off, n, blockIdx.x);
    const EwVlArgs& a = A[r];
    const int id = blockIdx.x - off[r];
    if (id >= a.nblk) return;
    ewald_vl_body<float, true>(a.srt, a.srf, a.perm, a.cnt, a.nex, a.nl, a.na, a.nb, a.cap, a.L0, a.L1, a.L2, a.i0,
                               a.i1, a.i2, a.rc2, a.alpha, a.sfac, a.grad, a.parts, a.dmax, a.tab, a.tsc, a.tku, a.tkr,
                               id);
}
__global__ __launch_bounds__(512) void k_sk_h16_batch(const SkH16Args* __restrict__ A, const int* __restrict__ off,
                                                      int n) {
    const int r = batch_find(off, n, blockIdx.x);
    const SkH16Args& a = A[r];
    const int id = blockIdx.x - off[r];
    if (id >= a.nblk) return;
    sk_h16_body(a.ph, a.q, a.na, a.chunk, a.nchunk, a.kd, a.skp, id % a.nchunk, id / a.nchunk);
}
__global__ __launch_bounds__(256) void k_sk_final_batch(const SkFinArgs* __restrict__ A, const int* __restrict__ off,
                                                        int n) {
    const int r = batch_find(off, n, blockIdx.x);
    const SkFinArgs& a = A[r];
    const int id = blockIdx.x - off[r];
    if (id >= a.nblk) return;
    const int nkx = a.kd.kmax + 1;
    sk_final_body<float2, true>(a.skp, a.nchunk, a.kd, a.mask, a.gk, a.gks, a.parts, id % nkx, id / nkx);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_kforce_h16_batch(
    const KfH16Args* __restrict__ A, const int* __restrict__ off, int n) {
    const int r = batch_find(off, n, blockIdx.x);
    const KfH16Args& a = A[r];
    const int id = blockIdx.x - off[r];
    if (id >= a.nblk) return;
    kforce_h16_body(a.ph, a.q, a.na, a.nbk, a.kd, a.ga, a.gks, a.kpart, id);
}

struct ForceBatch {
    int n = 0;
    // what was uploaded (records zero-filled, so that they compare bytewise)
    std::vector<EwVlArgs> ew;
    std::vector<SkH16Args> sk;
    std::vector<SkFinArgs> sf;
    std::vector<KfH16Args> kf;
    std::vector<int> off;  // [4][n + 1] range starts: pairs, k_sk_h16, k_sk_final, k_kforce_h16
    char* d = nullptr;
    size_t cap = 0, o_ew = 0, o_sk = 0, o_sf = 0, o_kf = 0, o_off = 0;
};

ForceBatch* force_batch_new() { return new ForceBatch(); }
void force_batch_free(ForceBatch* fb) {
    if (!fb) return;
    if (fb->d) (void)hipFree(fb->d);
    delete fb;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int force_batch_set(ForceBatch* fb, const FBItem* it, int n, hipStream_t s, bool capturing) {
    std::vector<EwVlArgs> ew(n);
    std::vector<SkH16Args> sk(n);
    std::vector<SkFinArgs> sf(n);
    std::vector<KfH16Args> kf(n);
    std::vector<int> off((size_t)4 * (n + 1), 0);
    memset(ew.data(), 0, sizeof(EwVlArgs) * n);
    memset(sk.data(), 0, sizeof(SkH16Args) * n);
    memset(sf.data(), 0, sizeof(SkFinArgs) * n);
    memset(kf.data(), 0, sizeof(KfH16Args) * n);
    for (int i = 0; i < n; i++) {
        const FBItem& x = it[i];
        const VList& v = *x.vl;
        const int na = x.na, nb = x.nb;
        EwVlArgs& e = ew[i];
        e.srt = v.srt;
        e.srf = v.srf;
        e.perm = v.perm;
        e.cnt = v.cnt;
        e.nex = v.nex;
        e.nl = v.nl;
        e.na = na;
        e.nb = nb;
        e.cap = v.cap;
        e.nblk = ewald_real_parts(na) * nb;
        e.L0 = x.sb.L[0];
        e.L1 = x.sb.L[1];
        e.L2 = x.sb.L[2];
        e.i0 = x.sb.inv[0];
        e.i1 = x.sb.inv[1];
        e.i2 = x.sb.inv[2];
        e.rc2 = x.rcut * x.rcut;
        e.alpha = x.alpha;
        e.sfac = -2.0 * (x.alpha / sqrt(acos(-1.0)));  // as launch_ewald_real
        e.grad = x.grad;
        e.parts = x.parts;
        e.dmax = v.dmax;
        e.tab = x.tab->d;
        e.tsc = x.tab->c;
        e.tku = x.tab->ku;
        e.tkr = x.tab->kr;
        const KDev kd = make_kdev(*x.ks);
        int chunk = 0;
        const int nchunk = kspace_nchunk_h16(na, nb, &chunk);
        SkH16Args& a = sk[i];
        a.ph = (const double2*)x.ph;
        a.q = x.q;
        a.skp = (float2*)x.skp;
        a.na = na;
        a.chunk = chunk;
        a.nchunk = nchunk;
        a.nblk = nchunk * nb;
        a.kd = kd;
        SkFinArgs& f = sf[i];
        f.skp = (const float2*)x.skp;
        f.mask = x.kmask;
        f.gk = (double2*)x.gk;
        f.gks = x.gks;
        f.parts = x.parts_k;
        f.nchunk = nchunk;
        f.nblk = (x.ks->kmax + 1) * nb;
        f.kd = kd;
        KfH16Args& k = kf[i];
        k.ph = (const double2*)x.ph;
        k.q = x.q;
        k.ga = (const uint4*)x.gk;
        k.gks = x.gks;
        k.kpart = x.kpart;
        k.na = na;
        k.nbk = nb;
        k.nblk = (na + KFH - 1) / KFH * ((nb + 7) / 8 * 8);
        k.kd = kd;
        const int nblk[4] = {e.nblk, a.nblk, f.nblk, k.nblk};
        for (int t = 0; t < 4; t++)
            off[(size_t)t * (n + 1) + i + 1] = off[(size_t)t * (n + 1) + i] + (nblk[t] + 7) / 8 * 8;
    }
    const bool same = fb->n == n && fb->d && !memcmp(ew.data(), fb->ew.data(), sizeof(EwVlArgs) * n) &&
                      !memcmp(sk.data(), fb->sk.data(), sizeof(SkH16Args) * n) &&
                      !memcmp(sf.data(), fb->sf.data(), sizeof(SkFinArgs) * n) &&
                      !memcmp(kf.data(), fb->kf.data(), sizeof(KfH16Args) * n) && off == fb->off;
    if (same) return 0;
    if (capturing) return 1;
    size_t o = 0;
    const size_t o_ew = o;
    o = align256(o + sizeof(EwVlArgs) * n);
    const size_t o_sk = o;
    o = align256(o + sizeof(SkH16Args) * n);
    const size_t o_sf = o;
    o = align256(o + sizeof(SkFinArgs) * n);
    const size_t o_kf = o;
    o = align256(o + sizeof(KfH16Args) * n);
    const size_t o_off = o;
    o = align256(o + sizeof(int) * off.size());
    // queued launches may still read the old records
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (o > fb->cap) {
        if (fb->d) (void)hipFree(fb->d);
        fb->d = nullptr;
        fb->cap = 0;
        if (hipMalloc((void**)&fb->d, o) != hipSuccess) return -1;
        fb->cap = o;
    }
    std::vector<char> h(o, 0);
    memcpy(h.data() + o_ew, ew.data(), sizeof(EwVlArgs) * n);
    memcpy(h.data() + o_sk, sk.data(), sizeof(SkH16Args) * n);
    memcpy(h.data() + o_sf, sf.data(), sizeof(SkFinArgs) * n);
    memcpy(h.data() + o_kf, kf.data(), sizeof(KfH16Args) * n);
    memcpy(h.data() + o_off, off.data(), sizeof(int) * off.size());
    if (hipMemcpyAsync(fb->d, h.data(), o, hipMemcpyHostToDevice, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return -1;
    fb->n = n;
    fb->ew.swap(ew);
    fb->sk.swap(sk);
    fb->sf.swap(sf);
    fb->kf.swap(kf);
    fb->off.swap(off);
    fb->o_ew = o_ew;
    fb->o_sk = o_sk;
    fb->o_sf = o_sf;
    fb->o_kf = o_kf;
    fb->o_off = o_off;
    return 0;
}

// the four launches: pairs on sr, then (or beside them, sr != sk) the k-space chain on sk.  bpc > 0
// caps the pair kernel at bpc blocks per CU (dynamic LDS).
void force_batch_launch(const ForceBatch* fb, hipStream_t sk, hipStream_t sr, int bpc) {
    const int n = fb->n;
    const int* off = (const int*)(fb->d + fb->o_off);
    auto total = [&](int t) { return (unsigned)fb->off[(size_t)t * (n + 1) + n]; };
    static size_t dyn[9] = {~(size_t)0, ~(size_t)0, ~(size_t)0, ~(size_t)0, ~(size_t)0,
                            ~(size_t)0, ~(size_t)0, ~(size_t)0, ~(size_t)0};  // [cap], caps 0 .. 8
    size_t& d = dyn[std::max(0, std::min(bpc, 8))];
    if (d == ~(size_t)0) d = lds_for_blocks_per_cu((const void*)k_ewald_vl32_batch, std::min(bpc, 8));
    k_ewald_vl32_batch<<<total(0), 256, d, sr>>>((const EwVlArgs*)(fb->d + fb->o_ew), off, n);
    k_sk_h16_batch<<<total(1), 512, 0, sk>>>((const SkH16Args*)(fb->d + fb->o_sk), off + (n + 1), n);
    k_sk_final_batch<<<total(2), 256, 0, sk>>>((const SkFinArgs*)(fb->d + fb->o_sf), off + 2 * (n + 1), n);
    k_kforce_h16_batch<<<total(3), 256, 0, sk>>>((const KfH16Args*)(fb->d + fb->o_kf), off + 3 * (n + 1), n);
}

// ------------------------------------------------------------------------------------------------
// O-O neighbour lists on centroids (src/neigh_list.f90; gating src/ljones.f90:144-163) and the
// LJ / Buckingham pair sums (src/ljones.f90, src/buckingham.f90); TEM correction
// (src/correct_tem.f90).  Lists are full (both directions), capacity RPMD_NCAP per atom.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void mimage(double& dx, double& dy, double& dz, const Box& bx) {
    dx -= bx.L[0] * rint(dx * bx.inv[0]);
    dy -= bx.L[1] * rint(dy * bx.inv[1]);
    dz -= bx.L[2] * rint(dz * bx.inv[2]);
}

// one wave per listed atom i (stride njump); cen is SoA [3][na].  Lanes test 64 candidates at a
// time; ballot + prefix count keep the list in ascending j, so it is deterministic.
__global__ __launch_bounds__(256) void k_nlist(const double* __restrict__ cen, int na, int njump, Box bx, double cut2,
                                               int* __restrict__ nnb, int* __restrict__ nbl, int* __restrict__ overflow) {
    const int n = na / njump;
    const int ii = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (ii >= n) return;
    const int i = ii * njump;
    const double xi = cen[i], yi = cen[na + i], zi = cen[2 * na + i];
    int cnt = 0;
    for (int base = 0; base < n; base += 64) {
        const int jj = base + lane;
        bool in = false;
        int j = 0;
        if (jj < n) {
            j = jj * njump;
            if (j != i) {
                double dx = cen[j] - xi, dy = cen[na + j] - yi, dz = cen[2 * na + j] - zi;
                mimage(dx, dy, dz, bx);
                in = dx * dx + dy * dy + dz * dz < cut2;
            }
        }
        const unsigned long long m = __ballot(in);
        if (in) {
            const int slot = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (slot < RPMD_NCAP) nbl[(size_t)ii * RPMD_NCAP + slot] = j;
        }
        cnt += __popcll(m);
    }
    if (lane == 0) {
        if (cnt > RPMD_NCAP) atomicMax(overflow, cnt);
        if (cnt > 0) atomicAdd(overflow + 1, cnt);
        nnb[ii] = cnt < RPMD_NCAP ? cnt : RPMD_NCAP;
    }
}
void launch_nlist(const double* cen, int na, int njump, Box box, double cut, int* nnb, int* nbl, int* overflow,
                  hipStream_t s) {
    const int n = na / njump;
    k_nlist<<<(n + 3) / 4, 256, 0, s>>>(cen, na, njump, box, cut * cut, nnb, nbl, overflow);
}

// O-O centroid lists of the device-resident calls, one wave per oxygen.  When the gate is on, the
// Verlet list of the oxygens whose centroids lie within rl = cut + skin (ascending, as k_nlist) and
// cb = the oxygen centroids of this build; the list stays valid while no oxygen centroid has moved
// more than skin/2.  Every call: the pairs within cut, taken from the Verlet list (or directly,
// when building) in its order and with the test of k_nlist, so the result is the list of k_nlist.
// vnnb/vnbl: the Verlet list; nnb/nbl: the list of the call; overflow[0] is sticky.  dmax is
// cleared by k_oo, after every wave has read the gate.
__global__ __launch_bounds__(256) void k_nlist_vl(const double* __restrict__ cen, int na, Box bx, double rl2,
                                                  double cut2, int* __restrict__ vnnb, int* __restrict__ vnbl,
                                                  int* __restrict__ nnb, int* __restrict__ nbl,
                                                  double* __restrict__ cb, int* __restrict__ overflow, VGate gate) {
    const int n = na / 3;
    const int ii = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (ii >= n) return;
    const int i = 3 * ii;
    const double xi = cen[i], yi = cen[na + i], zi = cen[2 * na + i];
    int cnt = 0;
    if (gate.on()) {
        if (lane == 0) {
            cb[ii] = xi;
            cb[n + ii] = yi;
            cb[2 * n + ii] = zi;
        }
        int cv = 0;
        for (int base = 0; base < n; base += 64) {
            const int jj = base + lane;
            bool inv = false, inc = false;
            int j = 0;
            if (jj < n) {
                j = 3 * jj;
                if (j != i) {
                    double dx = cen[j] - xi, dy = cen[na + j] - yi, dz = cen[2 * na + j] - zi;
                    mimage(dx, dy, dz, bx);
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    inv = r2 < rl2;
                    inc = r2 < cut2;
                }
            }
            const unsigned long long mv = __ballot(inv), mc = __ballot(inc);
            const unsigned long long below = (1ull << lane) - 1ull;
            if (inv) {
                const int slot = cv + __popcll(mv & below);
                if (slot < RPMD_NCAP) vnbl[(size_t)ii * RPMD_NCAP + slot] = j;
            }
            if (inc) {
                const int slot = cnt + __popcll(mc & below);
                if (slot < RPMD_NCAP) nbl[(size_t)ii * RPMD_NCAP + slot] = j;
            }
            cv += __popcll(mv);
            cnt += __popcll(mc);
        }
        if (lane == 0) {
            if (cv > RPMD_NCAP) overflow[0] = 1;
            vnnb[ii] = cv < RPMD_NCAP ? cv : RPMD_NCAP;
        }
    } else {
        const int cv = vnnb[ii];
        for (int base = 0; base < cv; base += 64) {
            const int e = base + lane;
            bool in = false;
            int j = 0;
            if (e < cv) {
                j = vnbl[(size_t)ii * RPMD_NCAP + e];
                double dx = cen[j] - xi, dy = cen[na + j] - yi, dz = cen[2 * na + j] - zi;
                mimage(dx, dy, dz, bx);
                in = dx * dx + dy * dy + dz * dz < cut2;
            }
            const unsigned long long m = __ballot(in);
            if (in) nbl[(size_t)ii * RPMD_NCAP + cnt + __popcll(m & ((1ull << lane) - 1ull))] = j;
            cnt += __popcll(m);
        }
    }
    if (lane == 0) nnb[ii] = cnt < RPMD_NCAP ? cnt : RPMD_NCAP;
}
// gate: an oxygen centroid coordinate moved by more than skin / (2 sqrt 3)
void launch_nlist_vl(const double* cen, int na, Box box, double cut, double skin, int force, int* vnnb, int* vnbl,
                     double* cb, unsigned long long* dmax, int* overflow, int* nnb, int* nbl, hipStream_t s) {
    const VGate gate{dmax, 0.25 * skin * skin / 3.0, force};
    const double rl = cut + skin;
    k_nlist_vl<<<(unsigned)((na / 3 + 3) / 4), 256, 0, s>>>(cen, na, box, rl * rl, cut * cut, vnnb, vnbl, nnb, nbl, cb,
                                                           overflow, gate);
}

// LJ (gam == 0) or Buckingham exp-6 (src/ljones.f90, src/buckingham.f90); OOL lanes per (bead, O),
// lane s taking neighbours s, s + OOL, ...; the O gradient is summed over its lanes by a fixed
// butterfly.  use_list = 0: every other O (the reference's basic / cell branches, same pair set).
#define OOL 8
// ch.parts (optional): the Ewald chain rule (k_chain) of the molecule is applied here as well, its
// three lanes writing grad, with the O-O gradient added to the oxygen; dmax (optional) is cleared.
struct ChainArgs {
    const double* parts = nullptr;
    int npart = 0;
    size_t pstride = 0;
    double alpha = 1.0, scale = 1.0;
    int acc = 0;
};
template <typename T>
__global__ __launch_bounds__(256) void k_oo(const double* __restrict__ pos, int na, Box bx, const int* __restrict__ nnb,
                                            const int* __restrict__ nbl, int use_list, double rc2, double sigsq,
                                            double ba, double bb, double bc, int buck, double vs, double ds,
                                            ChainArgs ch, unsigned long long* __restrict__ dmax,
                                            double* __restrict__ grad, double* __restrict__ parts,
                                            double* __restrict__ lj, int nb) {
    // blocks are dealt to the eight XCDs in turn (block id mod 8): XCD x takes beads x, x + 8, ... one
    // after the other, so the blocks resident on an XCD read one bead's positions from its L2
    const int n = na / 3, nbb = (n * OOL + 255) / 256, m8 = blockIdx.x >> 3;
    const int b = (blockIdx.x & 7) + 8 * (m8 / nbb), blk = m8 % nbb;
    if (b >= nb) return;  // the whole block
    const int t = blk * blockDim.x + threadIdx.x;
    const int ii = t / OOL, sl = t % OOL;
    if (dmax && t == 0 && b == 0) ctrl_store(dmax, 0ull);
    double acc[7];
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (ii < n) {
        const int i = 3 * ii;
        const double* X = pos + (size_t)b * 3 * na;
        const double xi = X[i], yi = X[na + i], zi = X[2 * na + i];
        const int cnt = use_list ? nnb[ii] : n;
        // the lane's entries k = sl, sl + OOL, ... in groups of four, their loads issued together
        for (int k0 = sl; k0 < cnt; k0 += 4 * OOL) {
            int jv[4];
            bool ok[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u * OOL;
                ok[u] = k < cnt && (use_list || k != ii);
                jv[u] = !ok[u] ? i : (use_list ? nbl[(size_t)ii * RPMD_NCAP + k] : 3 * k);
            }
            double xj[4], yj[4], zj[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                xj[u] = X[jv[u]];
                yj[u] = X[na + jv[u]];
                zj[u] = X[2 * na + jv[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (!ok[u]) continue;
                double dx = xi - xj[u], dy = yi - yj[u], dz = zi - zj[u];
                mimage(dx, dy, dz, bx);
                const double dr2 = dx * dx + dy * dy + dz * dz;
                if (dr2 < rc2) {
                    double v, fr;   // fr: dV/dr / r (gradient on i = fr * (r_i - r_j))
                    if (!buck) {
                        const T onr2 = (T)1 / (T)dr2;
                        const T sr2 = (T)sigsq * onr2;
                        const T sr6 = sr2 * sr2 * sr2;
                        v = vs * (double)(sr6 * (sr6 - (T)1));
                        fr = -ds * (double)(sr6 * (sr6 - (T)0.5) * onr2);
                    } else {
                        const T tr = sqrt((T)dr2);
                        const T br6 = (T)bb / ((T)dr2 * (T)dr2 * (T)dr2);
                        const T ex = (T)ba * exp((T)bc * tr);
                        v = (double)(ex - br6);
                        fr = (double)(((T)bc * ex + (T)6 * br6 / tr) / tr);
                    }
                    const double fx = fr * dx, fy = fr * dy, fz = fr * dz;
                    gx += fx;
                    gy += fy;
                    gz += fz;
                    acc[0] += v;
                    acc[1] += dx * fx;
                    acc[2] += dx * fy;
                    acc[3] += dx * fz;
                    acc[4] += dy * fy;
                    acc[5] += dy * fz;
                    acc[6] += dz * fz;
                }
            }
        }
    }
#pragma unroll
    for (int off = 1; off < OOL; off <<= 1) {
        gx += __shfl_xor(gx, off, 64);
        gy += __shfl_xor(gy, off, 64);
        gz += __shfl_xor(gz, off, 64);
    }
    if (ii < n && !ch.parts && sl == 0) {
        if (lj) {
            // stored per molecule for the chain-rule kernel to add
            lj[((size_t)b * 3 + 0) * n + ii] = gx;
            lj[((size_t)b * 3 + 1) * n + ii] = gy;
            lj[((size_t)b * 3 + 2) * n + ii] = gz;
        } else {
            const int i = 3 * ii;
            double* G = grad + (size_t)b * 3 * na;
            G[i] += gx;
            G[na + i] += gy;
            G[2 * na + i] += gz;
        }
    }
    if (ii < n && ch.parts && sl < 3) {
        // k_chain for dimension sl of the molecule, then the oxygen's O-O gradient
        const int d = sl;
        const size_t o = ((size_t)b * 3 + d) * na + 3 * ii;
        double s0 = ch.parts[o], s1 = ch.parts[o + 1], s2 = ch.parts[o + 2];
        for (int p = 1; p < ch.npart; p++) {
            s0 += ch.parts[p * ch.pstride + o];
            s1 += ch.parts[p * ch.pstride + o + 1];
            s2 += ch.parts[p * ch.pstride + o + 2];
        }
        const double alpha2 = 0.5 * (1.0 - ch.alpha);
        const double g0 = ch.scale * s0, g1 = ch.scale * s1, g2 = ch.scale * s2;
        const double b0 = ch.acc ? grad[o] : 0.0, b1 = ch.acc ? grad[o + 1] : 0.0, b2 = ch.acc ? grad[o + 2] : 0.0;
        const double lj = d == 0 ? gx : (d == 1 ? gy : gz);
        grad[o] = (b0 + ch.alpha * g0) + lj;
        grad[o + 1] = b1 + (g1 + alpha2 * g0);
        grad[o + 2] = b2 + (g2 + alpha2 * g0);
    }
    // every pair is seen from both ends: energy and virial carry 1/2; the virial of a central
    // force is symmetric
    double w[10] = {0.5 * acc[0], 0.5 * acc[1], 0.5 * acc[2], 0.5 * acc[3], 0.5 * acc[2],
                    0.5 * acc[4], 0.5 * acc[5], 0.5 * acc[3], 0.5 * acc[5], 0.5 * acc[6]};
    block_reduce_store<10>(w, parts + ((size_t)b * nbb + blk) * 10);
}

template <typename T>
void launch_oo(const double* pos, int na, int nb, Box box, const int* nnb, const int* nbl, int use_list, double rcut,
               const rpmd_params& p, double* grad, double* parts, hipStream_t s, const double* chain_parts,
               int chain_npart, double chain_alpha, double chain_scale, int chain_acc, unsigned long long* dmax,
               double* lj) {
    ChainArgs ch;
    ch.parts = chain_parts;
    ch.npart = chain_npart;
    ch.pstride = (size_t)3 * na * nb;
    ch.alpha = chain_alpha;
    ch.scale = chain_scale;
    ch.acc = chain_acc;
    const int buck = p.oo_gam != 0.0;
    double ba = 0, bb = 0, bc = 0;
    if (buck) {
        ba = exp(p.oo_gam) * 6.0 * p.oo_eps / (p.oo_gam - 6.0);
        bb = p.oo_eps * pow(p.oo_sig, 6) / (1.0 - (6.0 / p.oo_gam));
        bc = -p.oo_gam / p.oo_sig;
    }
    const int n = na / 3;
    const double vs = buck ? 1.0 : 4.0 * p.oo_eps, ds = buck ? 1.0 : 48.0 * p.oo_eps;
    k_oo<T><<<oo_blocks(na) * ((nb + 7) / 8 * 8), 256, 0, s>>>(pos, na, box, nnb, nbl, use_list, rcut * rcut,
                                                                p.oo_sig * p.oo_sig, ba, bb, bc, buck, vs, ds, ch,
                                                                dmax, grad, parts, lj, nb);
}
template void launch_oo<double>(const double*, int, int, Box, const int*, const int*, int, double, const rpmd_params&,
                                double*, double*, hipStream_t, const double*, int, double, double, int,
                                unsigned long long*, double*);
template void launch_oo<float>(const double*, int, int, Box, const int*, const int*, int, double, const rpmd_params&,
                               double*, double*, hipStream_t, const double*, int, double, double, int,
                               unsigned long long*, double*);
int oo_blocks(int na) { return (na / 3 * OOL + 255) / 256; }

// TEM switched short-range Coulomb between different molecules (src/correct_tem.f90:73-177);
// site holds physical M-site positions; thread per (bead, atom)
__global__ __launch_bounds__(256) void k_tem(const double* __restrict__ site, const double* __restrict__ q, int na, Box bx,
                                             const int* __restrict__ nnb, const int* __restrict__ nbl, int use_list,
                                             double sig, double* __restrict__ gsite, double* __restrict__ parts) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    if (i < na) {
        const double* X = site + (size_t)b * 3 * na;
        const double xi = X[i], yi = X[na + i], zi = X[2 * na + i], qi = q[i];
        const double rc2 = sig * sig, a1 = 2.0 / sig, a2 = 2.0 / (sig * sig * sig), a3 = 1.0 / (sig * sig * sig * sig);
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const int cnt = use_list ? nnb[i] : na;
        for (int k = 0; k < cnt; k++) {
            const int j = use_list ? nbl[(size_t)i * RPMD_NCAP + k] : k;
            if (j / 3 == i / 3) continue;
            double dx = xi - X[j], dy = yi - X[na + j], dz = zi - X[2 * na + j];
            mimage(dx, dy, dz, bx);
            const double drsq = dx * dx + dy * dy + dz * dz;
            if (drsq < rc2) {
                const double dr = sqrt(drsq), onr = 1.0 / dr, zij = qi * q[j];
                const double vc = zij * onr, dvc = -vc * onr, dr3 = drsq * dr, dr4 = dr3 * dr;
                const double f = 1.0 - a1 * dr + a2 * dr3 - a3 * dr4;
                const double df = -a1 + 3.0 * a2 * drsq - 4.0 * a3 * dr3;
                const double dv = vc * f, dvr = (vc * df + dvc * f) * onr;
                const double fx = dvr * dx, fy = dvr * dy, fz = dvr * dz;
                gx += fx;
                gy += fy;
                gz += fz;
                acc[0] += 0.5 * dv;
                acc[1] += 0.5 * dx * fx; acc[2] += 0.5 * dy * fx; acc[3] += 0.5 * dz * fx;
                acc[4] += 0.5 * dx * fy; acc[5] += 0.5 * dy * fy; acc[6] += 0.5 * dz * fy;
                acc[7] += 0.5 * dx * fz; acc[8] += 0.5 * dy * fz; acc[9] += 0.5 * dz * fz;
            }
        }
        double* G = gsite + (size_t)b * 3 * na;
        G[i] = gx;
        G[na + i] = gy;
        G[2 * na + i] = gz;
    }
    block_reduce_store<10>(acc, parts + ((size_t)b * gridDim.x + blockIdx.x) * 10);
}
void launch_tem(const double* site, const double* q, int na, int nb, Box box, const int* nnb, const int* nbl,
                int use_list, double sig, double* gsite, double* parts, hipStream_t s) {
    k_tem<<<dim3((na + 255) / 256, nb), 256, 0, s>>>(site, q, na, box, nnb, nbl, use_list, sig, gsite, parts);
}
int tem_blocks(int na) { return (na + 255) / 256; }

// ------------------------------------------------------------------------------------------------
// intramolecular terms (src/intra_harmonic.f90, intra_spcfw.f90, intra_morse_harm.f90,
// intra_morse.f90); thread per (bead, molecule)
// ------------------------------------------------------------------------------------------------
struct IntraDev {
    int model;
    double apot, bpot, alp, alpb, theta, reoh;
};

__global__ __launch_bounds__(256) void k_intra(const double* __restrict__ pos, int na, IntraDev p,
                                               double* __restrict__ grad, double* __restrict__ parts, int accg) {
    const int b = blockIdx.y;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    if (m < na / 3) {
        const double* X = pos + (size_t)b * 3 * na;
        const int j = 3 * m;
        double dx[3][3], dr[3], u[3][3], sq[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            dx[0][c] = X[c * na + j + 1] - X[c * na + j];
            dx[1][c] = X[c * na + j + 2] - X[c * na + j];
            dx[2][c] = X[c * na + j + 2] - X[c * na + j + 1];
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sq[k] = dx[k][0] * dx[k][0] + dx[k][1] * dx[k][1] + dx[k][2] * dx[k][2];
            dr[k] = sqrt(sq[k]);
#pragma unroll
            for (int c = 0; c < 3; c++) u[k][c] = dx[k][c] / dr[k];
        }
        double v = 0.0, dvb[3];
        if (p.model == RPMD_INTRA_HARMONIC) {
            const double ap = 0.5993, bp = 0.1466, cp = -0.09435, dp = 0.04984;
            const double rehh = 2.0 * sin(0.5 * p.theta) * p.reoh;
            const double d1 = dr[0] - p.reoh, d2 = dr[1] - p.reoh, d3 = dr[2] - rehh;
            v = 0.5 * ap * (d1 * d1 + d2 * d2) + 0.5 * bp * d3 * d3 + cp * (d1 + d2) * d3 + dp * d1 * d2;
            dvb[0] = ap * d1 + cp * d3 + dp * d2;
            dvb[1] = ap * d2 + cp * d3 + dp * d1;
            dvb[2] = bp * d3 + cp * (d1 + d2);
        } else {
            const double uu = sq[0] + sq[1] - sq[2], vv = 2.0 * dr[0] * dr[1], v2 = 1.0 / (vv * vv);
            const double arg = uu / vv, darg = -1.0 / sqrt(1.0 - arg * arg), ang = acos(arg);
            const double dth0 = darg * ((2.0 * dr[0] * vv - 2.0 * dr[1] * uu) * v2);
            const double dth1 = darg * ((2.0 * dr[1] * vv - 2.0 * dr[0] * uu) * v2);
            const double dth2 = darg * ((-2.0 * dr[2] * vv) * v2);
            const double dang = ang - p.theta, d1 = dr[0] - p.reoh, d2 = dr[1] - p.reoh;
            if (p.model == RPMD_INTRA_SPCFW) {
                v = 0.5 * p.apot * (d1 * d1 + d2 * d2) + 0.5 * p.bpot * dang * dang;
                dvb[0] = p.apot * d1 + p.bpot * dang * dth0;
                dvb[1] = p.apot * d2 + p.bpot * dang * dth1;
                dvb[2] = p.bpot * dang * dth2;
            } else {
                const double de = p.apot, deb = p.bpot, al = p.alp, al2 = al * al, al3 = al * al2, al4 = al3 * al;
                const double f1 = 7.0 / 12.0, f2 = 7.0 / 3.0, ra = d1 * d1, rb = d2 * d2;
                v = de * (al2 * ra - al3 * d1 * ra + f1 * al4 * ra * ra) + de * (al2 * rb - al3 * d2 * rb + f1 * al4 * rb * rb);
                const double a1 = de * (2.0 * al2 * d1 - 3.0 * al3 * ra + f2 * al4 * d1 * ra);
                const double a2 = de * (2.0 * al2 * d2 - 3.0 * al3 * rb + f2 * al4 * d2 * rb);
                double a3;
                if (p.model == RPMD_INTRA_MORSE_HARM) {
                    v += deb * dang * dang;
                    a3 = 2.0 * deb * dang;
                } else {
                    const double ab = p.alpb, ab2 = ab * ab, ab3 = ab2 * ab, ab4 = ab2 * ab2;
                    v += deb * (ab2 * dang * dang - ab3 * dang * dang * dang + f1 * ab4 * dang * dang * dang * dang);
                    a3 = deb * (2.0 * ab2 * dang - 3.0 * ab3 * dang * dang + f2 * ab4 * dang * dang * dang);
                }
                dvb[0] = a1 + a3 * dth0;
                dvb[1] = a2 + a3 * dth1;
                dvb[2] = a3 * dth2;
            }
        }
        double dv[3][3];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int c = 0; c < 3; c++) dv[k][c] = dvb[k] * u[k][c];
        double* G = grad + (size_t)b * 3 * na;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t o = (size_t)c * na + j;
            G[o] = (accg ? G[o] : 0.0) + (-dv[0][c] - dv[1][c]);
            G[o + 1] = (accg ? G[o + 1] : 0.0) + (dv[0][c] - dv[2][c]);
            G[o + 2] = (accg ? G[o + 2] : 0.0) + (dv[1][c] + dv[2][c]);
        }
        acc[0] = v;
        if (p.model == RPMD_INTRA_SPCFW) {
            // W = sum r_i . dvdr_i, diagonal only (src/intra_spcfw.f90:215-220)
#pragma unroll
            for (int c = 0; c < 3; c++)
                acc[1 + c * 4] = X[c * na + j] * (-dv[0][c] - dv[1][c]) + X[c * na + j + 1] * (dv[0][c] - dv[2][c]) +
                                 X[c * na + j + 2] * (dv[1][c] + dv[2][c]);
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int a = 0; a < 3; a++)
                    acc[1 + c * 3 + a] = dx[0][a] * dv[0][c] + dx[1][a] * dv[1][c] + dx[2][a] * dv[2][c];
        }
    }
    block_reduce_store<10>(acc, parts + ((size_t)b * gridDim.x + blockIdx.x) * 10);
}

void launch_intra(const double* pos, int na, int nb, int model, const rpmd_params& p, double* grad, double* parts,
                  int acc, hipStream_t s) {
    IntraDev d{model, p.apot, p.bpot, p.alp, p.alpb, p.theta, p.reoh};
    k_intra<<<dim3((na / 3 + 255) / 256, nb), 256, 0, s>>>(pos, na, d, grad, parts, acc);
}
int intra_blocks(int na) { return (na / 3 + 255) / 256; }

}  // namespace rpmd
