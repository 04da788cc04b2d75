// k_step.hip -- propagation kernels: bead normal-mode transform, momentum kicks, free ring
// polymer, PILE-L kick with centroid COM removal, counter-based Gaussian noise.
//
// All arrays are SoA per bead: x[(k*3 + d)*na + a].  Element e = d*na + a of bead/mode k sits at
// x[k*3*na + e], so the bead transform is out[k][e] = sum_j M[k][j] in[j][e]: a dense nb x nb
// matrix applied to 3*na columns -- HBM bound (2*nb*8 bytes per column per call).
#include <algorithm>

#include "common.hpp"

namespace rpmd {

// ---- NM transform (src/fourier.f90:16-63 as a matrix) ------------------------------------------
template <int NBM>
__global__ __launch_bounds__(256) void k_nm_reg(const double* __restrict__ in, double* __restrict__ out,
                                                const double* __restrict__ M, int n, size_t ne) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double x[NBM];
#pragma unroll
    for (int j = 0; j < NBM; j++) x[j] = j < n ? in[(size_t)j * ne + e] : 0.0;
    for (int k = 0; k < n; k++) {
        const double* row = M + (size_t)k * n;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NBM; j++)
            if (j < n) s = fma(row[j], x[j], s);
        out[(size_t)k * ne + e] = s;
    }
}

__global__ __launch_bounds__(256) void k_nm_gen(const double* __restrict__ in, double* __restrict__ out,
                                                const double* __restrict__ M, int n, size_t ne) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    for (int k = 0; k < n; k++) {
        const double* row = M + (size_t)k * n;
        double s = 0.0;
        for (int j = 0; j < n; j++) s = fma(row[j], in[(size_t)j * ne + e], s);
        out[(size_t)k * ne + e] = s;
    }
}

// FP64 MFMA form (v_mfma_f64_16x16x4f64): D[i][jj] = sum_kk A[i][kk] B[kk][jj] with
// A = M (16 output modes x 4 input beads), B = in (4 input beads x 16 columns).  Lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; it receives D[(l/16) + 4v][l%16], v = 0..3.  One wave owns a
// 16-column tile; all NT*16 output rows of the tile accumulate over ceil(n/4) bead steps.
typedef double dbl4 __attribute__((ext_vector_type(4)));
template <int NT>
__global__ __launch_bounds__(256) void k_nm_mfma(const double* __restrict__ in, double* __restrict__ out,
                                                 const double* __restrict__ M, int n, size_t ne) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const size_t ntile = (ne + 15) / 16;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t nwave = ((size_t)gridDim.x * blockDim.x) >> 6;
    const int nstep = (n + 3) / 4;
    for (size_t t = wave; t < ntile; t += nwave) {
        const size_t e = t * 16 + li;
        const bool ein = e < ne;
        dbl4 acc[NT];
#pragma unroll
        for (int q = 0; q < NT; q++) acc[q] = (dbl4){0.0, 0.0, 0.0, 0.0};
        for (int st = 0; st < nstep; st++) {
            const int j = st * 4 + lk;
            const double b = (ein && j < n) ? in[(size_t)j * ne + e] : 0.0;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                const int k = q * 16 + li;
                const double a = (k < n && j < n) ? M[(size_t)k * n + j] : 0.0;
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
        }
        if (ein) {
#pragma unroll
            for (int q = 0; q < NT; q++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = q * 16 + lk + 4 * v;
                    if (k < n) out[(size_t)k * ne + e] = acc[q][v];
                }
        }
    }
}

// out may not alias in
void launch_nm(const double* in, double* out, const double* M, int n, size_t ne, hipStream_t s) {
    const unsigned g = (unsigned)((ne + 255) / 256);
    const size_t ntile = (ne + 15) / 16;
    const unsigned gm = (unsigned)std::min<size_t>((ntile + 3) / 4, 2048);  // 4 waves per block
    if (n <= 8) k_nm_reg<8><<<g, 256, 0, s>>>(in, out, M, n, ne);
    else if (n <= 16) k_nm_mfma<1><<<gm, 256, 0, s>>>(in, out, M, n, ne);
    else if (n <= 32) k_nm_mfma<2><<<gm, 256, 0, s>>>(in, out, M, n, ne);
    else if (n <= 48) k_nm_mfma<3><<<gm, 256, 0, s>>>(in, out, M, n, ne);
    else if (n <= 64) k_nm_mfma<4><<<gm, 256, 0, s>>>(in, out, M, n, ne);
    else k_nm_gen<<<g, 256, 0, s>>>(in, out, M, n, ne);
}

// ---- momentum kicks: p -= a1 g1 + a2 g2 (src/evolve_pi.f90:82-83, 158, 228, 279) ---------------
// Each term is applied as its own rounded update, as the reference does.
__global__ void k_kick(double* __restrict__ p, const double* __restrict__ g1, double a1, const double* __restrict__ g2,
                       double a2, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = p[i];
    if (g1) v = __dadd_rn(v, -__dmul_rn(a1, g1[i]));
    if (g2) v = __dadd_rn(v, -__dmul_rn(a2, g2[i]));
    p[i] = v;
}
void launch_kick(double* p, const double* g1, double a1, const double* g2, double a2, size_t n, hipStream_t s) {
    k_kick<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, g1, a1, g2, a2, n);
}

// ---- free ring polymer: monodromy per mode (src/evolve_pi.f90:109-125), no contraction -----------
// monod(3,4,nb) column-major as ring_rpmd fills it (src/md_pathi.f90:190-231)
__global__ void k_free_ring(double* __restrict__ p, double* __restrict__ r, const double* __restrict__ monod, int na,
                            int nb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ne = (size_t)3 * na;
    if (i >= ne * nb) return;
    const int k = (int)(i / ne);
    const int a = (int)((i % ne) % na);
    const int ic = a % 3;
    const double* mo = monod + (size_t)k * 12;
    const double pv = p[i], rv = r[i];
    const double pp = __dadd_rn(__dmul_rn(pv, mo[0 * 3 + ic]), __dmul_rn(rv, mo[1 * 3 + ic]));
    const double rr = __dadd_rn(__dmul_rn(pv, mo[2 * 3 + ic]), __dmul_rn(rv, mo[3 * 3 + ic]));
    p[i] = pp;
    r[i] = rr;
}
void launch_free_ring(double* p, double* r, const double* monod, int na, int nb, hipStream_t s) {
    const size_t n = (size_t)3 * na * nb;
    k_free_ring<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, r, monod, na, nb);
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

// ---- PILE-L kick in the NM representation (src/module_gle.f90:8-73) ------------------------------
// p(i,j,k) = c1(k) p + c2(k) sqrt(m(ic)) xi ; xi from the host buffer in the reference draw order
// (mode, atom, xyz) or from the counter-based stream.  Block partial sums of the centroid momentum
// go to com_parts[blk][3].
__global__ __launch_bounds__(256) void k_pile(double* __restrict__ p, const double* __restrict__ c1,
                                              const double* __restrict__ c3, int na, int nb,
                                              const double* __restrict__ noise, uint64_t seed,
                                              const uint64_t* __restrict__ streamp, uint64_t stream_add,
                                              double* __restrict__ com_parts) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const uint64_t stream = (streamp ? *streamp : 0) + stream_add;
    double s[3] = {0.0, 0.0, 0.0};
    if (a < na) {
        const int ic = a % 3;
        const double cc1 = c1[k], cc3 = c3[k * 3 + ic];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = ((size_t)k * 3 + d) * na + a;
            const size_t ni = ((size_t)k * na + a) * 3 + d;
            const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
            const double v = __dadd_rn(__dmul_rn(cc1, p[o]), __dmul_rn(cc3, xi));
            p[o] = v;
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
            com_parts[blockIdx.x * 3 + threadIdx.x] = t;
        }
    }
}
int pile_blocks(int na) { return (na + 255) / 256; }
void launch_pile(double* p, const double* c1, const double* c3, int na, int nb, const double* noise, uint64_t seed,
                 const uint64_t* stream, uint64_t stream_add, double* com_parts, hipStream_t s) {
    k_pile<<<dim3(pile_blocks(na), nb), 256, 0, s>>>(p, c1, c3, na, nb, noise, seed, stream, stream_add, com_parts);
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
    const uint64_t stream = (streamp ? *streamp : 0) + stream_add;
    if (i >= (size_t)na * nb) return;
    const size_t k = i / na, a = i % na;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const size_t o = (k * 3 + d) * na + a, ni = (k * na + a) * 3 + d;
        const double xi = noise ? noise[ni] : normal_at(seed, stream, ni);
        p[o] = __dadd_rn(__dmul_rn(c1, p[o]), __dmul_rn(c3m[a], xi));
    }
}
void launch_langevin(double* p, const double* c3m, double c1, int na, int nb, const double* noise, uint64_t seed,
                     const uint64_t* stream, uint64_t stream_add, hipStream_t s) {
    const size_t n = (size_t)na * nb;
    k_langevin<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, c3m, c1, na, nb, noise, seed, stream, stream_add);
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
__global__ void k_counter_inc(uint64_t* c) { *c += 16; }
void launch_counter_inc(uint64_t* c, hipStream_t s) { k_counter_inc<<<1, 1, 0, s>>>(c); }

}  // namespace rpmd
