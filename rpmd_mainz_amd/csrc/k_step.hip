// k_step.hip -- propagation kernels: bead normal-mode transform, momentum kicks, free ring
// polymer, PILE-L kick with centroid COM removal, counter-based Gaussian noise.
//
// All arrays are SoA per bead: x[(k*3 + d)*na + a].  Element e = d*na + a of bead/mode k sits at
// x[k*3*na + e], so the bead transform is out[k][e] = sum_j M[k][j] in[j][e]: a dense nb x nb
// matrix applied to 3*na columns -- HBM bound (2*nb*8 bytes per column per call).
#include <algorithm>

#include "common.hpp"

// The propagation and thermostat kernels round every product and sum once, as the reference's Fortran
// loops do when compiled without contraction: HIP's __dmul_rn and __dadd_rn are plain * and + here,
// which the compiler would otherwise fuse into FMAs.
#pragma clang fp contract(off)

namespace rpmd {

// ---- NM transform (src/fourier.f90:16-63 as a matrix) ------------------------------------------
// One launch transforms up to two arrays (NMSrc a, b) with the same matrix; b.in == nullptr: one.  An
// input element may be formed on load: in - a1 g1 - a2 g2 (each term rounded, as k_kick), and for
// mode 0 minus m_a pcom_d, pcom = (sum of com[nblk][3]) / M (as k_com_remove).
struct NMSrc {
    const double* in = nullptr;
    double* out = nullptr;
    const double* g1 = nullptr;
    double a1 = 0.0;
    const double* g2 = nullptr;
    double a2 = 0.0;
    const double* com = nullptr;
    int nblk = 0;
    double inv_mtot = 0.0;
    const double* mass = nullptr;
    int na = 0;
};
__device__ __forceinline__ double nm_load(const NMSrc& s, int j, size_t e, size_t ne) {
    const size_t o = (size_t)j * ne + e;
    double v = s.in[o];
    if (s.g1) v = rn_add(v, -rn_mul(s.a1, s.g1[o]));
    if (s.g2) v = rn_add(v, -rn_mul(s.a2, s.g2[o]));
    if (s.com && j == 0) {
        const int d = (int)(e / s.na), a = (int)(e % s.na);
        double pc = 0.0;
        for (int q = 0; q < s.nblk; q++) pc += s.com[q * 3 + d];
        const double pcd = pc * s.inv_mtot;
        v -= pcd * s.mass[a];
    }
    return v;
}

template <int NBM>
__global__ __launch_bounds__(256) void k_nm_reg(NMSrc sa, NMSrc sb, const double* __restrict__ M, int n, size_t ne) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const NMSrc& src = e < ne ? sa : sb;
    if (e >= ne) e -= ne;
    if (e >= ne || !src.in) return;
    double x[NBM];
#pragma unroll
    for (int j = 0; j < NBM; j++) x[j] = j < n ? nm_load(src, j, e, ne) : 0.0;
    for (int k = 0; k < n; k++) {
        const double* row = M + (size_t)k * n;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NBM; j++)
            if (j < n) s = fma(row[j], x[j], s);
        src.out[(size_t)k * ne + e] = s;
    }
}

__global__ __launch_bounds__(256) void k_nm_gen(NMSrc sa, NMSrc sb, const double* __restrict__ M, int n, size_t ne) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const NMSrc& src = e < ne ? sa : sb;
    if (e >= ne) e -= ne;
    if (e >= ne || !src.in) return;
    for (int k = 0; k < n; k++) {
        const double* row = M + (size_t)k * n;
        double s = 0.0;
        for (int j = 0; j < n; j++) s = fma(row[j], nm_load(src, j, e, ne), s);
        src.out[(size_t)k * ne + e] = s;
    }
}

// FP64 MFMA form (v_mfma_f64_16x16x4f64): D[i][jj] = sum_kk A[i][kk] B[kk][jj] with
// A = M (16 output modes x 4 input beads), B = in (4 input beads x 16 columns).  Lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; it receives D[(l/16) + 4v][l%16], v = 0..3.  One wave owns a
// 16-column tile of one source; all NT*16 output rows of the tile accumulate over ceil(n/4) bead
// steps.
typedef double dbl4 __attribute__((ext_vector_type(4)));
template <int NT>
__global__ __launch_bounds__(256) void k_nm_mfma(NMSrc sa, NMSrc sb, const double* __restrict__ M, int n, size_t ne) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const size_t ntile = (ne + 15) / 16, ntot = sb.in ? 2 * ntile : ntile;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t nwave = ((size_t)gridDim.x * blockDim.x) >> 6;
    const int nstep = (n + 3) / 4;
    for (size_t tt = wave; tt < ntot; tt += nwave) {
        const bool second = tt >= ntile;
        const NMSrc& src = second ? sb : sa;
        const size_t t = second ? tt - ntile : tt;
        const size_t e = t * 16 + li;
        const bool ein = e < ne;
        dbl4 acc[NT];
#pragma unroll
        for (int q = 0; q < NT; q++) acc[q] = (dbl4){0.0, 0.0, 0.0, 0.0};
        // the tile's inputs (NT*4 bead steps cover n <= 16 NT) are loaded before the MFMA chain,
        // so the loads are in flight together
        double bv[NT * 4];
#pragma unroll
        for (int st = 0; st < NT * 4; st++) {
            const int j = st * 4 + lk;
            bv[st] = (ein && j < n) ? nm_load(src, j, e, ne) : 0.0;
        }
#pragma unroll
        for (int st = 0; st < NT * 4; st++) {
            if (st >= nstep) break;
            const int j = st * 4 + lk;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const int k = q * 16 + li;
                const double a = (k < n && j < n) ? M[(size_t)k * n + j] : 0.0;
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[st], acc[q], 0, 0, 0);
            }
        }
        if (ein) {
#pragma unroll
            for (int q = 0; q < NT; q++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = q * 16 + lk + 4 * v;
                    if (k < n) src.out[(size_t)k * ne + e] = acc[q][v];
                }
        }
    }
}

// outputs may not alias inputs
static void launch_nm_src(const NMSrc& sa, const NMSrc& sb, const double* M, int n, size_t ne, hipStream_t s) {
    const size_t cols = sb.in ? 2 * ne : ne;
    const unsigned g = (unsigned)((cols + 255) / 256);
    const size_t ntile = (ne + 15) / 16 * (sb.in ? 2 : 1);
    const unsigned gm = (unsigned)std::min<size_t>((ntile + 3) / 4, 4096);  // 4 waves per block
    if (n <= 8) k_nm_reg<8><<<g, 256, 0, s>>>(sa, sb, M, n, ne);
    else if (n <= 16) k_nm_mfma<1><<<gm, 256, 0, s>>>(sa, sb, M, n, ne);
    else if (n <= 32) k_nm_mfma<2><<<gm, 256, 0, s>>>(sa, sb, M, n, ne);
    else if (n <= 48) k_nm_mfma<3><<<gm, 256, 0, s>>>(sa, sb, M, n, ne);
    else if (n <= 64) k_nm_mfma<4><<<gm, 256, 0, s>>>(sa, sb, M, n, ne);
    else k_nm_gen<<<g, 256, 0, s>>>(sa, sb, M, n, ne);
}
void launch_nm(const double* in, double* out, const double* M, int n, size_t ne, hipStream_t s) {
    NMSrc a;
    a.in = in;
    a.out = out;
    launch_nm_src(a, NMSrc{}, M, n, ne, s);
}
// pnm = F (p - a1 g1 - a2 g2) and rnm = F r in one launch (the kicked p itself is not stored)
void launch_nm_kick2(const double* p, const double* g1, double a1, const double* g2, double a2, double* pnm,
                     const double* r, double* rnm, const double* F, int n, size_t ne, hipStream_t s) {
    NMSrc a, b;
    a.in = p;
    a.out = pnm;
    a.g1 = g1;
    a.a1 = a1;
    a.g2 = g2;
    a.a2 = a2;
    b.in = r;
    b.out = rnm;
    launch_nm_src(a, b, F, n, ne, s);
}
// rst = B rnm and p = B (pnm with the centroid COM removed, as k_com_remove) in one launch
void launch_nm_com2(const double* rnm, double* rst, const double* pnm, double* p, const double* com, int nblk,
                    double inv_mtot, const double* mass, int na, const double* B, int n, size_t ne, hipStream_t s) {
    NMSrc a, b;
    a.in = rnm;
    a.out = rst;
    b.in = pnm;
    b.out = p;
    b.com = com;
    b.nblk = nblk;
    b.inv_mtot = inv_mtot;
    b.mass = mass;
    b.na = na;
    launch_nm_src(a, b, B, n, ne, s);
}

// ---- momentum kicks: p -= a1 g1 + a2 g2 (src/evolve_pi.f90:82-83, 158, 228, 279) ---------------
// Each term is applied as its own rounded update, as the reference does.
// ctr (optional): advanced by the noise streams of one step (nothing after this kick draws in the step)
__global__ void k_kick(double* __restrict__ p, const double* __restrict__ g1, double a1, const double* __restrict__ g2,
                       double a2, size_t n, uint64_t* __restrict__ ctr) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ctr && i == 0) ctrl_add(ctr, (uint64_t)RPMD_STREAMS_PER_STEP);
    if (i >= n) return;
    double v = p[i];
    if (g1) v = rn_add(v, -rn_mul(a1, g1[i]));
    if (g2) v = rn_add(v, -rn_mul(a2, g2[i]));
    p[i] = v;
}
void launch_kick(double* p, const double* g1, double a1, const double* g2, double a2, size_t n, hipStream_t s,
                 uint64_t* ctr) {
    k_kick<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, g1, a1, g2, a2, n, ctr);
}

// ---- free ring polymer: monodromy per mode (src/evolve_pi.f90:109-125), no contraction -----------
// monod(3,4,nb) column-major as ring_rpmd fills it (src/md_pathi.f90:190-231)
// com (optional): the centroid COM momentum is removed from mode 0 first (as k_com_remove)
__global__ void k_free_ring(double* __restrict__ p, double* __restrict__ r, const double* __restrict__ monod, int na,
                            int nb, const double* __restrict__ com, int nblk, double inv_mtot,
                            const double* __restrict__ mass) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ne = (size_t)3 * na;
    if (i >= ne * nb) return;
    const int k = (int)(i / ne);
    const int a = (int)((i % ne) % na);
    const int ic = a % 3;
    const double* mo = monod + (size_t)k * 12;
    double pv = p[i];
    const double rv = r[i];
    if (com && k == 0) {
        const int d = (int)(i / na);
        double pc = 0.0;
        for (int q = 0; q < nblk; q++) pc += com[q * 3 + d];
        const double pcd = pc * inv_mtot;
        pv -= pcd * mass[a];
    }
    const double pp = rn_add(rn_mul(pv, mo[0 * 3 + ic]), rn_mul(rv, mo[1 * 3 + ic]));
    const double rr = rn_add(rn_mul(pv, mo[2 * 3 + ic]), rn_mul(rv, mo[3 * 3 + ic]));
    p[i] = pp;
    r[i] = rr;
}
void launch_free_ring(double* p, double* r, const double* monod, int na, int nb, hipStream_t s, const double* com,
                      int nblk, double inv_mtot, const double* mass) {
    const size_t n = (size_t)3 * na * nb;
    k_free_ring<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, r, monod, na, nb, com, nblk, inv_mtot, mass);
}

// ---- counter-based Gaussian noise: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller ------------
__device__ __forceinline__ void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                             uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
    uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
}
__device__ __forceinline__ void philox4(uint64_t key, uint64_t ctr_hi, uint64_t ctr_lo, uint32_t out[4]) {
    uint32_t c0 = (uint32_t)ctr_lo, c1 = (uint32_t)(ctr_lo >> 32), c2 = (uint32_t)ctr_hi, c3 = (uint32_t)(ctr_hi >> 32);
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// standard normal for (stream, index); uses 53-bit uniforms
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint32_t o[4];
    philox4(seed, stream, idx, o);
    const double u1 = ((double)((((uint64_t)o[0]) << 21) ^ (o[1] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((double)((((uint64_t)o[2]) << 21) ^ (o[3] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// ---- PILE kick in the NM representation (src/module_gle.f90:8-73; modes k0.. for PILE-G :97-115) ----
// p(i,j,k) = c1(k) p + c2(k) sqrt(m(ic)) xi for modes k >= k0; xi from the host buffer in the
// reference draw order (mode, atom, xyz) or from the counter-based stream.  With com_parts, block
// partial sums of the centroid momentum go to com_parts[blk][3].
__global__ __launch_bounds__(256) void k_pile(double* __restrict__ p, const double* __restrict__ c1,
                                              const double* __restrict__ c3, int na, int k0,
                                              const double* __restrict__ noise, uint64_t seed,
                                              const uint64_t* __restrict__ streamp, uint64_t stream_add,
                                              double* __restrict__ com_parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y + k0;
    const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
    double s[3] = {0.0, 0.0, 0.0};
    if (a < na) {
        const int ic = a % 3;
        const double cc1 = c1[k], cc3 = c3[k * 3 + ic];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = ((size_t)k * 3 + d) * na + a;
            const size_t ni = ((size_t)(k - k0) * na + a) * 3 + d;
            const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
            const double v = rn_add(rn_mul(cc1, p[o]), rn_mul(cc3, xi));
            p[o] = v;
            s[d] = v;
        }
    }
    if (k == 0 && com_parts) {
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) s[d] += __shfl_down(s[d], off, 64);
        __shared__ double red[4][3];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0)
            for (int d = 0; d < 3; d++) red[w][d] = s[d];
        __syncthreads();
        if (threadIdx.x < 3) {
            double t = 0.0;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q][threadIdx.x];
            com_parts[blockIdx.x * 3 + threadIdx.x] = t;
        }
    }
}
int pile_blocks(int na) { return (na + 255) / 256; }

// The free ring step (as k_free_ring, the centroid COM in com_in removed from mode 0 first) and then
// the PILE kick (as k_pile, modes 0..nb-1) in one launch, thread per (atom, mode) as k_pile, so the
// partial centroid sums of the kick go to com_out in k_pile's layout and order
__global__ __launch_bounds__(256) void k_free_ring_pile(double* __restrict__ p, double* __restrict__ r,
                                                        const double* __restrict__ monod, int na,
                                                        const double* __restrict__ com_in, int nblk, double inv_mtot,
                                                        const double* __restrict__ mass,
                                                        const double* __restrict__ c1,
                                                        const double* __restrict__ c3,
                                                        const double* __restrict__ noise, uint64_t seed,
                                                        const uint64_t* __restrict__ streamp, uint64_t stream_add,
                                                        double* __restrict__ com_out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
    double s[3] = {0.0, 0.0, 0.0};
    if (a < na) {
        const int ic = a % 3;
        const double* mo = monod + (size_t)k * 12;
        const double cc1 = c1[k], cc3 = c3[k * 3 + ic];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t i = ((size_t)k * 3 + d) * na + a;
            double pv = p[i];
            const double rv = r[i];
            if (k == 0) {
                double pc = 0.0;
                for (int q = 0; q < nblk; q++) pc += com_in[q * 3 + d];
                const double pcd = pc * inv_mtot;
                pv -= pcd * mass[a];
            }
            const double pp = rn_add(rn_mul(pv, mo[0 * 3 + ic]), rn_mul(rv, mo[1 * 3 + ic]));
            const double rr = rn_add(rn_mul(pv, mo[2 * 3 + ic]), rn_mul(rv, mo[3 * 3 + ic]));
            r[i] = rr;
            const size_t ni = ((size_t)k * na + a) * 3 + d;
            const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
            const double v = rn_add(rn_mul(cc1, pp), rn_mul(cc3, xi));
            p[i] = v;
            s[d] = v;
        }
    }
    if (k == 0) {
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) s[d] += __shfl_down(s[d], off, 64);
        __shared__ double red[4][3];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0)
            for (int d = 0; d < 3; d++) red[w][d] = s[d];
        __syncthreads();
        if (threadIdx.x < 3) {
            double t = 0.0;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q][threadIdx.x];
            com_out[blockIdx.x * 3 + threadIdx.x] = t;
        }
    }
}
void launch_free_ring_pile(double* p, double* r, const double* monod, int na, int nb, const double* com_in,
                           double inv_mtot, const double* mass, const double* c1, const double* c3,
                           const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add,
                           double* com_out, hipStream_t s) {
    k_free_ring_pile<<<dim3(pile_blocks(na), nb), 256, 0, s>>>(p, r, monod, na, com_in, pile_blocks(na), inv_mtot,
                                                               mass, c1, c3, noise, seed, stream, stream_add,
                                                               com_out);
}
// ---- the PILE-L step up to the forces in two launches ------------------------------------------
// both deviates of one Box-Muller draw: z0 = normal_at(seed, stream, idx) and its sine partner
__device__ __forceinline__ void normal2_at(uint64_t seed, uint64_t stream, uint64_t idx, double& z0, double& z1) {
    uint32_t o[4];
    philox4(seed, stream, idx, o);
    const double u1 = ((double)((((uint64_t)o[0]) << 21) ^ (o[1] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((double)((((uint64_t)o[2]) << 21) ^ (o[3] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
    const double rad = sqrt(-2.0 * log(u1));
    z0 = rad * cospi(2.0 * u2);
    z1 = rad * sinpi(2.0 * u2);
}

// k_pile_step: one block of 2 NT waves per 16 columns e = d na + a of p and r (NT*16 >= nb modes).
// Wave (src, q) forms modes 16q.. of p (src 0, with the opening kick on load: p - a1 g1 - a2 g2, each
// term rounded, as k_kick) or r (src 1) by the forward transform (F, FP64 MFMA as k_nm_mfma) into
// LDS.  Then every thread takes two (mode, column) entries: the first PILE kick (as k_pile), the free
// ring step (as k_free_ring) and the second PILE kick; on the device stream the two kicks draw the
// cosine and sine deviates of one Box-Muller draw of the first kick's stream.  Then wave (src, q)
// forms beads 16q.. of p or rst by the inverse transform (B).  The centroid COM removals of the
// separate sequence (after each kick, src/module_gle.f90:55-70) are left out here: they are linear in
// the removed momentum, the first only shifts the centroid positions by mo_2 dp1 (its change of p is
// removed again by the second) and the second subtracts m_a P2 / M from the centroid momentum, P1
// and P2 being the centroid momentum sums after the first and the second kick as computed here.
// Block partial sums {P1, P2} go to parts[blk][6] and k_pile_com applies both corrections.
template <int NT>
__global__ __launch_bounds__(128 * NT) void k_pile_step(const double* __restrict__ p_in, const double* __restrict__ g1,
                                                        double a1, const double* __restrict__ g2, double a2,
                                                        const double* __restrict__ r_in, double* __restrict__ p_out,
                                                        double* __restrict__ r_out, const double* __restrict__ F,
                                                        const double* __restrict__ B, int n, int na,
                                                        const double* __restrict__ monod,
                                                        const double* __restrict__ c1,
                                                        const double* __restrict__ c3,
                                                        const double* __restrict__ noise1,
                                                        const double* __restrict__ noise2, uint64_t seed,
                                                        const uint64_t* __restrict__ streamp, uint64_t sadd1,
                                                        double* __restrict__ parts) {
    constexpr int NW = 2 * NT, KM = 16 * NT;
    __shared__ double T[2][KM][16];
    __shared__ double red[NW][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int src = w / NT, q = w % NT;
    const size_t ne = (size_t)3 * na, e0 = (size_t)blockIdx.x * 16;
    // every load below is unconditional, from a clamped (valid) index, with the value selected
    // afterwards: a load under a branch would be waited for before the next one is issued
    const int nm1 = n - 1;
    const size_t ec = e0 + li < ne ? e0 + li : ne - 1;
    double fa[NT * 4], ba[NT * 4];
    {
        const int k = q * 16 + li, kc = min(k, nm1);
#pragma unroll
        for (int st = 0; st < NT * 4; st++) {
            const int j = st * 4 + lk, jc = min(j, nm1);
            const bool kj = k < n && j < n;
            const double f = F[(size_t)kc * n + jc], bm = B[(size_t)kc * n + jc];
            fa[st] = kj ? f : 0.0;
            ba[st] = kj ? bm : 0.0;
        }
    }
    {
        const bool ein = e0 + li < ne;
        const double* in = src == 0 ? p_in : r_in;
        double xin[NT * 4], xg1[NT * 4], xg2[NT * 4];
#pragma unroll
        for (int st = 0; st < NT * 4; st++) xin[st] = in[(size_t)min(st * 4 + lk, nm1) * ne + ec];
        if (src == 0) {
            const double* g2v = g2 ? g2 : g1;
#pragma unroll
            for (int st = 0; st < NT * 4; st++) {
                const size_t o = (size_t)min(st * 4 + lk, nm1) * ne + ec;
                xg1[st] = g1[o];
                xg2[st] = g2v[o];
            }
        }
        double bv[NT * 4];
#pragma unroll
        for (int st = 0; st < NT * 4; st++) {
            double v = xin[st];
            if (src == 0) {
                v = rn_add(v, -rn_mul(a1, xg1[st]));
                if (g2) v = rn_add(v, -rn_mul(a2, xg2[st]));
            }
            bv[st] = ein && st * 4 + lk < n ? v : 0.0;
        }
        dbl4 acc = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < NT * 4; st++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[st], bv[st], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) T[src][q * 16 + lk + 4 * v][li] = acc[v];
    }
    // this thread's two (mode, column) entries: coefficients and deviates, loaded and drawn before
    // the transform results are waited for
    const uint64_t s0 = streamp ? ctrl_load(streamp) : 0;
    double ec1[2], ec3[2], emo[2][4], ex1[2], ex2[2];
    int ed[2];
    bool eok[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int idx = threadIdx.x + h * 128 * NT, k = idx >> 4, col = idx & 15;
        const size_t e = e0 + col < ne ? e0 + col : ne - 1;
        eok[h] = e0 + col < ne && k < n;
        const int kc = min(k, nm1), d = (int)(e / na), a = (int)(e % na), ic = a % 3;
        ed[h] = d;
        ec1[h] = c1[kc];
        ec3[h] = c3[kc * 3 + ic];
#pragma unroll
        for (int u = 0; u < 4; u++) emo[h][u] = monod[(size_t)kc * 12 + u * 3 + ic];
        const size_t ni = ((size_t)kc * na + a) * 3 + d;
        if (noise1) {
            ex1[h] = noise1[ni];
            ex2[h] = noise2[ni];
        } else {
            normal2_at(seed, s0 + sadd1, ni, ex1[h], ex2[h]);
        }
    }
    __syncthreads();
    double sv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int idx = threadIdx.x + h * 128 * NT, k = idx >> 4, col = idx & 15;
        if (eok[h]) {
            const double pv = rn_add(rn_mul(ec1[h], T[0][k][col]), rn_mul(ec3[h], ex1[h]));
            const double rv = T[1][k][col];
            const double pp = rn_add(rn_mul(pv, emo[h][0]), rn_mul(rv, emo[h][1]));
            const double rr = rn_add(rn_mul(pv, emo[h][2]), rn_mul(rv, emo[h][3]));
            const double pk = rn_add(rn_mul(ec1[h], pp), rn_mul(ec3[h], ex2[h]));
            T[0][k][col] = pk;
            T[1][k][col] = rr;
            if (k == 0) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    sv[c] += c == ed[h] ? pv : 0.0;
                    sv[3 + c] += c == ed[h] ? pk : 0.0;
                }
            }
        } else {
            T[0][k][col] = 0.0;
            T[1][k][col] = 0.0;
        }
    }
    __syncthreads();
    {
        const size_t e = e0 + li;
        dbl4 acc = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < NT * 4; st++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ba[st], T[src][st * 4 + lk][li], acc, 0, 0, 0);
        if (e < ne) {
            double* out = src == 0 ? p_out : r_out;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int kb = q * 16 + lk + 4 * v;
                if (kb < n) out[(size_t)kb * ne + e] = acc[v];
            }
        }
    }
    // block partial sums {P1, P2} of the centroid momentum, fixed order
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 6; c++) sv[c] += __shfl_down(sv[c], off, 64);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 6; c++) red[w][c] = sv[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
        for (int u = 0; u < NW; u++) t += red[u][threadIdx.x];
        parts[blockIdx.x * 6 + threadIdx.x] = t;
    }
}

// the centroid COM corrections of k_pile_step, thread per element e of PS_B beads: every block sums
// the partials in the same fixed order into {P1, P2}, then rst += B[b][0] mo_2 (-(P1 / M) m_a) and
// p += B[b][0] (-(P2 / M) m_a)
#define PS_B 8
__global__ __launch_bounds__(256) void k_pile_com(double* __restrict__ p, double* __restrict__ r,
                                                  const double* __restrict__ parts, int nblk,
                                                  const double* __restrict__ B, int n, int na,
                                                  const double* __restrict__ monod, const double* __restrict__ mass,
                                                  double inv_mtot) {
    // all loads first, unconditional from clamped indices (a load under a branch is waited for
    // before the next): this element's values of PS_B beads, then the partials
    const size_t ne = (size_t)3 * na;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ein = e < ne;
    const size_t ec = ein ? e : ne - 1;
    const int d = (int)(ec / na), a = (int)(ec % na), ic = a % 3;
    const double ma = mass[a], mo2 = monod[2 * 3 + ic];
    const int b0 = blockIdx.y * PS_B;
    double rv[PS_B], pv[PS_B], cb[PS_B];
#pragma unroll
    for (int i = 0; i < PS_B; i++) {
        const int bc = min(b0 + i, n - 1);
        const size_t o = (size_t)bc * ne + ec;
        rv[i] = r[o];
        pv[i] = p[o];
        cb[i] = B[(size_t)bc * n];
    }
    double sv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < nblk; q0 += 4 * (int)blockDim.x) {
        double pa[4][6];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int q = min(q0 + (int)threadIdx.x + u * (int)blockDim.x, nblk - 1);
#pragma unroll
            for (int c = 0; c < 6; c++) pa[u][c] = parts[q * 6 + c];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (q0 + (int)threadIdx.x + u * (int)blockDim.x < nblk)
#pragma unroll
                for (int c = 0; c < 6; c++) sv[c] += pa[u][c];
    }
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 6; c++) sv[c] += __shfl_down(sv[c], off, 64);
    __shared__ double red[4][6];
    __shared__ double tot[6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 6; c++) red[w][c] = sv[c];
    __syncthreads();
    if (threadIdx.x < 6) tot[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    __syncthreads();
    if (!ein) return;
    const double dr = mo2 * (-(tot[d] * inv_mtot) * ma), dp = -(tot[3 + d] * inv_mtot) * ma;
#pragma unroll
    for (int i = 0; i < PS_B; i++)
        if (b0 + i < n) {
            const size_t o = (size_t)(b0 + i) * ne + e;
            r[o] = rv[i] + cb[i] * dr;
            p[o] = pv[i] + cb[i] * dp;
        }
}

int pile_step_blocks(int na) { return (int)(((size_t)3 * na + 15) / 16); }

int launch_pile_step(const double* p_in, const double* g1, double a1, const double* g2, double a2, const double* r_in,
                     double* p_out, double* r_out, const double* F, const double* B, int n, int na,
                     const double* monod, const double* mass, double inv_mtot, const double* c1, const double* c3,
                     const double* noise1, const double* noise2, uint64_t seed, const uint64_t* stream,
                     uint64_t sadd1, double* parts, hipStream_t s) {
    const int nblk = pile_step_blocks(na);
#define PSTEP(NT)                                                                                               \
    k_pile_step<NT><<<nblk, 128 * NT, 0, s>>>(p_in, g1, a1, g2, a2, r_in, p_out, r_out, F, B, n, na, monod, c1, c3, \
                                              noise1, noise2, seed, stream, sadd1, parts)
    if (n <= 16)
        PSTEP(1);
    else if (n <= 32)
        PSTEP(2);
    else if (n <= 48)
        PSTEP(3);
    else if (n <= 64)
        PSTEP(4);
    else
        return 0;
#undef PSTEP
    const unsigned ge = (unsigned)(((size_t)3 * na + 255) / 256);
    k_pile_com<<<dim3(ge, (n + PS_B - 1) / PS_B), 256, 0, s>>>(p_out, r_out, parts, nblk, B, n, na, monod, mass,
                                                              inv_mtot);
    return 1;
}

void launch_pile(double* p, const double* c1, const double* c3, int na, int nb, int k0, const double* noise,
                 uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* com_parts, hipStream_t s) {
    if (nb <= k0) return;
    k_pile<<<dim3(pile_blocks(na), nb - k0), 256, 0, s>>>(p, c1, c3, na, k0, noise, seed, stream, stream_add,
                                                         com_parts);
}

// p(:,j,1) -= pcom * m(j), pcom = sum_j p(:,j,1) / sum m  (src/module_gle.f90:55-70)
__global__ void k_com_remove(double* __restrict__ p, const double* __restrict__ mass, const double* __restrict__ parts,
                             int nblk, double inv_mtot, int na) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    double pc[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < nblk; q++)
#pragma unroll
        for (int d = 0; d < 3; d++) pc[d] += parts[q * 3 + d];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double pcd = pc[d] * inv_mtot;
        p[(size_t)d * na + a] -= pcd * mass[a];
    }
}
void launch_com_remove(double* p, const double* mass, const double* com_parts, int nblk, double inv_mtot, int na,
                       hipStream_t s) {
    k_com_remove<<<(na + 255) / 256, 256, 0, s>>>(p, mass, com_parts, nblk, inv_mtot, na);
}

// ---- local Langevin on beads, PRL (src/module_thermostat.f90:56-82) -----------------------------
__global__ void k_langevin(double* __restrict__ p, const double* __restrict__ c3m, double c1, int na, int nb,
                           const double* __restrict__ noise, uint64_t seed, const uint64_t* __restrict__ streamp,
                           uint64_t stream_add) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
    if (i >= (size_t)na * nb) return;
    const size_t k = i / na, a = i % na;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const size_t o = (k * 3 + d) * na + a, ni = (k * na + a) * 3 + d;
        const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
        p[o] = rn_add(rn_mul(c1, p[o]), rn_mul(c3m[a], xi));
    }
}
void launch_langevin(double* p, const double* c3m, double c1, int na, int nb, const double* noise, uint64_t seed,
                     const uint64_t* stream, uint64_t stream_add, hipStream_t s) {
    const size_t n = (size_t)na * nb;
    k_langevin<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, c3m, c1, na, nb, noise, seed, stream, stream_add);
}

// ---- Bussi global rescale (src/module_thermostat.f90:6-54) ---------------------------------------
// One draw block per call: [0] rt (drawn, unused by the reference), [1, ndof) the chi-square
// deviates, [ndof] ranno.  kt sums p^2/(2m) over the first nbq beads (modes) of p.  Three launches:
// fixed-grid partial sums, one block that forms alpha in device memory, and the rescale.
#define BUSSI_BLOCKS 256
__global__ __launch_bounds__(256) void k_bussi_parts(const double* __restrict__ p, const double* __restrict__ mass,
                                                     int na, int nbq, long long ndof,
                                                     const double* __restrict__ noise, uint64_t seed,
                                                     const uint64_t* __restrict__ streamp, uint64_t stream_add,
                                                     double* __restrict__ parts) {
    const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
    const size_t n = (size_t)3 * na * nbq;
    const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, nth = (size_t)gridDim.x * 256;
    double v[2] = {0.0, 0.0};
    for (size_t e = tid; e < n; e += nth) {
        const double x = p[e];
        v[0] += x * x / (2.0 * mass[e % na]);
    }
    for (long long i = 1 + (long long)tid; i < ndof; i += (long long)nth) {
        const double g = noise ? noise[i] : normal_at(seed, stream, (uint64_t)i);
        v[1] += g * g;
    }
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 2; k++) v[k] += __shfl_down(v[k], off, 64);
    __shared__ double red[4][2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[w][0] = v[0];
        red[w][1] = v[1];
    }
    __syncthreads();
    if (threadIdx.x < 2)
        parts[blockIdx.x * 2 + threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                              red[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_bussi_alpha(const double* __restrict__ parts, int nblk, long long ndof,
                                                     double betan, double c, const double* __restrict__ noise,
                                                     uint64_t seed, const uint64_t* __restrict__ streamp,
                                                     uint64_t stream_add, double* __restrict__ alpha) {
    __shared__ double red[256][2];
    double a0 = 0.0, a1 = 0.0;
    for (int q = threadIdx.x; q < nblk; q += 256) {
        a0 += parts[q * 2];
        a1 += parts[q * 2 + 1];
    }
    red[threadIdx.x][0] = a0;
    red[threadIdx.x][1] = a1;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[threadIdx.x][0] += red[threadIdx.x + h][0];
            red[threadIdx.x][1] += red[threadIdx.x + h][1];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
        const double kt = red[0][0], snfm1 = red[0][1];
        const double ranno = noise ? noise[ndof] : normal_at(seed, stream, (uint64_t)ndof);
        const double ransq = ranno * ranno;
        const double ktav = (double)ndof / (2.0 * betan);
        const double fac = (1.0 - c) * ktav / ((double)ndof * kt);
        const double t1 = fac * (snfm1 + ransq);
        const double t2 = 2 * ranno * sqrt(c * fac);
        double al = sqrt(c + t1 + t2);
        const double plusmin = ranno + sqrt(c / fac);
        al = plusmin >= 0.0 ? fabs(al) : -fabs(al);
        alpha[0] = al;
    }
}

// p *= alpha on the first nbq beads; with com_parts, partial centroid momenta as k_pile writes them
__global__ __launch_bounds__(256) void k_bussi_scale(double* __restrict__ p, const double* __restrict__ alpha, int na,
                                                     double* __restrict__ com_parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const double al = alpha[0];
    double s[3] = {0.0, 0.0, 0.0};
    if (a < na) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = ((size_t)k * 3 + d) * na + a;
            const double v = al * p[o];
            p[o] = v;
            s[d] = v;
        }
    }
    if (k == 0 && com_parts) {
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) s[d] += __shfl_down(s[d], off, 64);
        __shared__ double red[4][3];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0)
            for (int d = 0; d < 3; d++) red[w][d] = s[d];
        __syncthreads();
        if (threadIdx.x < 3) {
            double t = 0.0;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q][threadIdx.x];
            com_parts[blockIdx.x * 3 + threadIdx.x] = t;
        }
    }
}

int bussi_parts_size() { return 2 * BUSSI_BLOCKS; }
void launch_bussi(double* p, const double* mass, int na, int nbq, long long ndof, double betan, double c,
                  const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* parts,
                  double* alpha, double* com_parts, hipStream_t s) {
    k_bussi_parts<<<BUSSI_BLOCKS, 256, 0, s>>>(p, mass, na, nbq, ndof, noise, seed, stream, stream_add, parts);
    k_bussi_alpha<<<1, 256, 0, s>>>(parts, BUSSI_BLOCKS, ndof, betan, c, noise, seed, stream, stream_add, alpha);
    k_bussi_scale<<<dim3(pile_blocks(na), nbq), 256, 0, s>>>(p, alpha, na, com_parts);
}

// ---- GLE kick (src/module_gle.f90:229-283) -------------------------------------------------------
// Per element (bead b, xyz d, atom a) of the (3 na nb) x (ns+1) matrix gp: channel 0 <- p/sqrt(m);
// gp <- gp T^T + xi S^T (T = TS[0..], S = TS[64..], column-major N1 x N1); xi in the reference draw
// order (channel, bead, atom, xyz).  Block partial sums of sqrt(m) gp per channel and xyz feed the
// centre-of-mass removal of pshift_gle (:285-320).
template <int N1>
__global__ __launch_bounds__(256) void k_gle(const double* __restrict__ p, double* __restrict__ gp,
                                             const double* __restrict__ TS, const double* __restrict__ mass, int na,
                                             int nb, const double* __restrict__ noise, uint64_t seed,
                                             const uint64_t* __restrict__ streamp, uint64_t stream_add,
                                             double* __restrict__ parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const uint64_t stream = (streamp ? ctrl_load(streamp) : 0) + stream_add;
    const size_t N = (size_t)3 * na * nb;
    double pc[N1 * 3];
#pragma unroll
    for (int i = 0; i < N1 * 3; i++) pc[i] = 0.0;
    if (a < na) {
        const double m = mass[a];
        const double mfac = 1.0 / sqrt(m), sm = sqrt(m);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t e = ((size_t)b * 3 + d) * na + a;
            double g[N1], x[N1], o[N1];
            g[0] = p[e] * mfac;
#pragma unroll
            for (int i = 1; i < N1; i++) g[i] = gp[i * N + e];
#pragma unroll
            for (int i = 0; i < N1; i++) {
                const size_t ni = (size_t)i * N + ((size_t)b * na + a) * 3 + d;
                x[i] = noise ? noise[ni] : normal_at(seed, stream, ni);
            }
#pragma unroll
            for (int j = 0; j < N1; j++) {
                double t = 0.0;
#pragma unroll
                for (int c = 0; c < N1; c++) t += g[c] * TS[c * N1 + j];
                double u = 0.0;
#pragma unroll
                for (int c = 0; c < N1; c++) u += x[c] * TS[64 + c * N1 + j];
                o[j] = t + u;
            }
#pragma unroll
            for (int i = 0; i < N1; i++) {
                gp[i * N + e] = o[i];
                pc[i * 3 + d] = o[i] * sm;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < N1 * 3; i++) pc[i] += __shfl_down(pc[i], off, 64);
    __shared__ double red[4][N1 * 3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < N1 * 3; i++) red[w][i] = pc[i];
    __syncthreads();
    if ((int)threadIdx.x < N1 * 3) {
        double t = 0.0;
        for (int q = 0; q < 4; q++) t += red[q][threadIdx.x];
        parts[((size_t)b * gridDim.x + blockIdx.x) * (N1 * 3) + threadIdx.x] = t;
    }
}

// pcom[i][d] = (sum of the partials) * mm * mm, mm = 1/sqrt(nb sum m)  (fixed order)
__global__ void k_gle_com(const double* __restrict__ parts, int nparts, int nv, double mm, double* __restrict__ pcom) {
    const int t = threadIdx.x;
    if (t >= nv) return;
    double s = 0.0;
    for (int q = 0; q < nparts; q++) s += parts[(size_t)q * nv + t];
    pcom[t] = s * mm * mm;
}

// gp -= pcom sqrt(m) on every channel, then p = gp(channel 0) sqrt(m)
__global__ __launch_bounds__(256) void k_gle_shift(double* __restrict__ p, double* __restrict__ gp,
                                                   const double* __restrict__ pcom, const double* __restrict__ mass,
                                                   int na, int nb, int n1) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (a >= na) return;
    const size_t N = (size_t)3 * na * nb;
    const double sm = sqrt(mass[a]);
    for (int d = 0; d < 3; d++) {
        const size_t e = ((size_t)b * 3 + d) * na + a;
        for (int i = 0; i < n1; i++) gp[i * N + e] = gp[i * N + e] - pcom[i * 3 + d] * sm;
        p[e] = gp[e] * sm;
    }
}

int gle_parts_size(int na, int nb) { return pile_blocks(na) * nb * 24; }
void launch_gle(double* p, double* gp, const double* TS, const double* mass, int na, int nb, int n1,
                const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* parts,
                double* pcom, double mm, hipStream_t s) {
    const dim3 grid(pile_blocks(na), nb);
    switch (n1) {
#define GLE_CASE(K) \
    case K: k_gle<K><<<grid, 256, 0, s>>>(p, gp, TS, mass, na, nb, noise, seed, stream, stream_add, parts); break;
        GLE_CASE(1) GLE_CASE(2) GLE_CASE(3) GLE_CASE(4) GLE_CASE(5) GLE_CASE(6) GLE_CASE(7) GLE_CASE(8)
#undef GLE_CASE
        default: return;
    }
    k_gle_com<<<1, 64, 0, s>>>(parts, (int)(grid.x * grid.y), n1 * 3, mm, pcom);
    k_gle_shift<<<grid, 256, 0, s>>>(p, gp, pcom, mass, na, nb, n1);
}

// initial auxiliary momenta gp = chol(C) xi with C = (nb/beta) I (src/module_gle.f90:196-213)
__global__ void k_gle_draw(double* __restrict__ gp, double sc, int na, int nb, int n1, uint64_t seed, uint64_t stream) {
    const size_t N = (size_t)3 * na * nb;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * n1) return;
    gp[t] = sc * normal_at(seed, stream, t);
}
void launch_gle_draw(double* gp, double sc, int na, int nb, int n1, uint64_t seed, uint64_t stream, hipStream_t s) {
    const size_t n = (size_t)3 * na * nb * n1;
    k_gle_draw<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(gp, sc, na, nb, n1, seed, stream);
}

// ---- sample_rpmd (src/module_thermostat.f90:199-226) and pshift (:229-264) ----------------------
// p(l,j,k) = sqrt(m_j / beta_n) xi in the draw order (bead, atom, xyz); block partial sums of p per
// xyz for the centre-of-mass shift
__global__ __launch_bounds__(256) void k_sample(double* __restrict__ p, const double* __restrict__ mass, int na,
                                                double betan, const double* __restrict__ noise, uint64_t seed,
                                                uint64_t stream, double* __restrict__ parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    double s[3] = {0.0, 0.0, 0.0};
    if (a < na) {
        const double sigma = sqrt(mass[a] / betan);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t ni = ((size_t)k * na + a) * 3 + d;
            const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
            const double v = sigma * xi;
            p[((size_t)k * 3 + d) * na + a] = v;
            s[d] = v;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; d++) s[d] += __shfl_down(s[d], off, 64);
    __shared__ double red[4][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int d = 0; d < 3; d++) red[w][d] = s[d];
    __syncthreads();
    if (threadIdx.x < 3)
        parts[((size_t)k * gridDim.x + blockIdx.x) * 3 + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// v = (sum of the partials) / (nb sum m), in fixed order
__global__ void k_pshift_v(const double* __restrict__ parts, int nparts, double totmass, double* __restrict__ v) {
    if (threadIdx.x >= 3) return;
    double s = 0.0;
    for (int q = 0; q < nparts; q++) s += parts[(size_t)q * 3 + threadIdx.x];
    v[threadIdx.x] = s / totmass;
}

__global__ void k_pshift_apply(double* __restrict__ p, const double* __restrict__ mass, const double* __restrict__ v,
                               int na, int nb) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)na * nb) return;
    const size_t k = t / na, a = t % na;
#pragma unroll
    for (int d = 0; d < 3; d++) p[(k * 3 + d) * na + a] = p[(k * 3 + d) * na + a] - mass[a] * v[d];
}

void launch_sample_rpmd(double* p, const double* mass, int na, int nb, double beta, double totmass,
                        const double* noise, uint64_t seed, uint64_t stream, double* parts, double* v,
                        hipStream_t s) {
    const dim3 grid(pile_blocks(na), nb);
    k_sample<<<grid, 256, 0, s>>>(p, mass, na, beta / (double)nb, noise, seed, stream, parts);
    k_pshift_v<<<1, 64, 0, s>>>(parts, (int)(grid.x * grid.y), totmass, v);
    const size_t n = (size_t)na * nb;
    k_pshift_apply<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, mass, v, na, nb);
}

// ---- contraction copies (src/evolve_pi.f90:377-431): modes 0..h and their mirrors ---------------
// dst[k] = scale * src[map[k]] for k < ndst (map[k] < 0: zero)
__global__ void k_mode_copy(double* __restrict__ dst, const double* __restrict__ src, const int* __restrict__ map,
                            int ndst, size_t ne, double scale) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * ndst) return;
    const int k = (int)(i / ne);
    const size_t e = i % ne;
    const int m = map[k];
    dst[i] = m >= 0 ? scale * src[(size_t)m * ne + e] : 0.0;
}
void launch_mode_copy(double* dst, const double* src, const int* map, int ndst, int na, double scale, hipStream_t s) {
    const size_t ne = (size_t)3 * na, n = ne * ndst;
    k_mode_copy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dst, src, map, ndst, ne, scale);
}

__global__ void k_scale(double* y, double a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * y[i];
}
void launch_scale(double* y, double a, size_t n, hipStream_t s) {
    if (n) k_scale<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(y, a, n);
}

// per-step noise stream counter kept on the device so that a captured step replays with fresh noise
// ---- estimators of the state (src/kinetic.f90:1-54, 248-265; src/md_eq.f90:255-278) -------------
// Thread per atom over every bead and component: kinetic energy sum_k |p|^2 / (2 m); spring energy
// sum_k m wn^2 |r_k - r_{k-1}|^2 / 2 (k - 1 cyclic); and the centroid-virial sums
// sum_k (r_k - r_c) . dvdr_k per xyz, for the low- (g1) and high-frequency (g2) gradients.  Block
// partial sums in fixed order, parts[blk][10] = {tk, vrp, tq1 xyz, tq2 xyz, 0, 0}.
__global__ __launch_bounds__(256) void k_estimators(const double* __restrict__ p, const double* __restrict__ r,
                                                    const double* __restrict__ g1, const double* __restrict__ g2,
                                                    const double* __restrict__ mass, int na, int nb, double wn2,
                                                    double* __restrict__ parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (a < na) {
        const double m = mass[a];
        double p2 = 0.0, dist = 0.0;
        for (int d = 0; d < 3; d++) {
            double rc = 0.0;
            for (int k = 0; k < nb; k++) rc += r[((size_t)k * 3 + d) * na + a];
            rc = rc / (double)nb;
            for (int k = 0; k < nb; k++) {
                const size_t o = ((size_t)k * 3 + d) * na + a;
                const size_t op = ((size_t)(k == 0 ? nb - 1 : k - 1) * 3 + d) * na + a;
                const double pk = p[o], rk = r[o], dx = rk - r[op], dr = rk - rc;
                p2 += pk * pk;
                dist += dx * dx;
                v[2 + d] += dr * g1[o];
                v[5 + d] += dr * g2[o];
            }
        }
        v[0] = p2 / (2.0 * m);
        v[1] = 0.5 * m * wn2 * dist;
    }
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 10; c++) v[c] += __shfl_down(v[c], off, 64);
    __shared__ double red[4][10];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 10; c++) red[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 10)
        parts[(size_t)blockIdx.x * 10 + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}
int estimator_blocks(int na) { return (na + 255) / 256; }
void launch_estimators(const double* p, const double* r, const double* g1, const double* g2, const double* mass,
                       int na, int nb, double wn2, double* parts, hipStream_t s) {
    k_estimators<<<estimator_blocks(na), 256, 0, s>>>(p, r, g1, g2, mass, na, nb, wn2, parts);
}

// ---- barostat scaling of the molecules (src/module_barostat.f90:66-79, 107-120, 176-190) ---------
// Thread per (molecule j, axis d): rcm = the bead average of the SUM of the molecule's three atom
// coordinates (center_water, src/md_tools.f90:367-390: not their mean), each sum ((c + x_1) + x_2) +
// x_3 in the reference's order, times 1/nb; every atom of the molecule in every bead then moves by
// rcm (scale_d - 1).  Plain roundings (no fused multiply-add), as the reference's loops.
__global__ __launch_bounds__(256) void k_baro_shift(double* __restrict__ r, int na, int nb, double s0, double s1,
                                                    double s2) {
    const int nm = na / 3;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nm) return;
    const int d = t / nm, j = t % nm;
    const double sc = d == 0 ? s0 : (d == 1 ? s1 : s2);
    double c = 0.0;
    for (int k = 0; k < nb; k++) {
        const double* x = r + ((size_t)k * 3 + d) * na + 3 * j;
        c = __dadd_rn(__dadd_rn(__dadd_rn(c, x[0]), x[1]), x[2]);
    }
    const double sh = __dmul_rn(__dmul_rn(c, 1.0 / (double)nb), sc - 1.0);
    for (int k = 0; k < nb; k++) {
        double* x = r + ((size_t)k * 3 + d) * na + 3 * j;
        x[0] = __dadd_rn(x[0], sh);
        x[1] = __dadd_rn(x[1], sh);
        x[2] = __dadd_rn(x[2], sh);
    }
}
void launch_baro_shift(double* r, int na, int nb, const double scale[3], hipStream_t s) {
    const int n = 3 * (na / 3);
    if (n) k_baro_shift<<<(n + 255) / 256, 256, 0, s>>>(r, na, nb, scale[0], scale[1], scale[2]);
}

__global__ void k_counter_inc(uint64_t* c) { ctrl_add(c, (uint64_t)RPMD_STREAMS_PER_STEP); }
void launch_counter_inc(uint64_t* c, hipStream_t s) { k_counter_inc<<<1, 1, 0, s>>>(c); }

}  // namespace rpmd
