// common.hpp -- shared device/host definitions of the MI355X RPMD path.
//
// Device layout (all in HBM, one context per host thread):
//   positions, gradients, momenta : SoA per bead, X[b][d][a] = base[(b*3 + d)*na + a], FP64
//   Ewald sites                   : same SoA layout; the oxygen replaced by the M-site and the
//                                   coordinates scaled to the unit box when the box is cubic
//   per-block energy/virial parts : double[10] = {v, vir(3,3) column-major}
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "../../include/rpmd_mi355x.h"

#define RPMD_KTILE 32   // k-space tile: ky, kz in [-16, 15] (index k + 16) -> kmax <= 15
#define RPMD_KMAX_DEV 15
#define RPMD_KTILE_BIG 64  // kmax 16 .. 31: ky, kz in [-32, 31] (index k + 32), FP64 kernels
#define RPMD_KMAX_BIG 31
#define RPMD_NCAP 512   // O-O / TEM neighbour capacity per atom (full list)
#define RPMD_STREAMS_PER_STEP 64  // counter-based noise streams reserved per propagation step

namespace rpmd {

// Control words that one kernel of a step writes and a later kernel reads -- the list displacement
// gates (max-reduced, read, cleared), the rebuild count, the noise-stream counter -- are accessed
// with agent-scope atomics only: the accesses go to the device-coherent level, never through a
// per-CU cache or an XCD-local L2 line, whichever CUs and XCDs the kernels of a step land on
template <typename T>
__device__ __forceinline__ T ctrl_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void ctrl_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void ctrl_add(T* p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void ctrl_max(T* p, T v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One FP64 product or sum, rounded once: no fusion into an FMA with a neighbouring operation (HIP's
// __dmul_rn / __dadd_rn are plain * and +, which the compiler contracts).  The propagation kernels
// use these to reproduce the reference's loops compiled without contraction bit for bit.
__device__ __forceinline__ double rn_mul(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double rn_add(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}

struct Box {
    double L[3], inv[3];
};

struct CellGrid {
    int nc[3];  // cells per box side (at most 1024, the Morton code's 10 bits per axis)
    int ncell;  // nc[0]*nc[1]*nc[2]
};

struct KSet {
    int kmax;
    int ky_max[RPMD_KMAX_DEV + 1];    // per kx: largest |ky| with a member (-1: none)
    int kz_bound[RPMD_KMAX_DEV + 1];  // per kx: largest |kz| with a member
    signed char kz_row[RPMD_KMAX_DEV + 1][RPMD_KTILE];  // per (kx, ky + 16): largest |kz| with a member, -1: none
    double lat[3];                    // 2 pi (cubic, unit box) or 2 pi / L_d
    double f0;                        // 2 pi (cubic) or 2 pi / V
    double b;                         // 1/(4 alpha^2)
    int cubic;
};

// the k set of kmax 16 .. 31 on 64 x 64 tiles (same meaning as KSet's fields)
struct KSetBig {
    int kmax;
    int ky_max[RPMD_KMAX_BIG + 1];
    int kz_bound[RPMD_KMAX_BIG + 1];
    signed char kz_row[RPMD_KMAX_BIG + 1][RPMD_KTILE_BIG];
    double lat[3];
    double f0;
    double b;
    int cubic;
};

// ---- k_forces.hip -----------------------------------------------------------------------------
// energy/virial part ranges reduced by one launch: job j sums parts[j][p*10 + k] into out[out[j] + k]
#define RED_JOBS 5
struct ReduceJobs {
    const double* parts[RED_JOBS];
    int nparts[RED_JOBS];
    int out[RED_JOBS];
    double scale[RED_JOBS];
    int n = 0;
    void add(const double* p, int np, int o, double sc) {
        parts[n] = p;
        nparts[n] = np;
        out[n] = o;
        scale[n] = sc;
        n++;
    }
};
void launch_reduce_ev(const ReduceJobs& jobs, double* out, hipStream_t s);
void launch_aos_to_soa(const double* aos, double* soa, int na, int nb, hipStream_t s);
void launch_soa_to_aos(const double* soa, double* aos, int na, int nb, hipStream_t s);
void launch_sites(const double* pos, double* site, int na, int nb, double alpha, int unit, Box box, hipStream_t s);
// the closing momentum kick of a step, p = (p - a1 g1) - a2 grad, fused into the chain-rule kernel that
// writes grad (ctr: the noise counter it advances)
struct ReduceJobs;
struct KickArgs {
    double* p = nullptr;
    const double* g1 = nullptr;
    double a1 = 0.0, a2 = 0.0;
    uint64_t* ctr = nullptr;
    // (optional) energy reductions run by extra blocks of the same launch (launch_reduce_ev)
    const ReduceJobs* red = nullptr;
    double* ev_out = nullptr;
};
// lj (optional): the oxygens' O-O gradient per molecule, [bead][dim][molecule], added to the oxygens;
// kick (optional): then the kick of the atoms
void launch_chain_rule(double* grad, const double* parts, int npart, int na, int nb, double alpha, double scale,
                       int acc, hipStream_t s, const double* lj = nullptr, const KickArgs* kick = nullptr);
void launch_centroid(const double* pos, double* cen, int na, int nb, hipStream_t s);
void launch_centroid_vl(const double* pos, double* cen, int na, int nb, const double* cb, unsigned long long* dmax,
                        hipStream_t s);
void launch_axpy(double* y, const double* x, double a, size_t n, hipStream_t s);
void launch_zero(double* y, size_t n, hipStream_t s);
void launch_zero_int(int* y, size_t n, hipStream_t s);
// every 32-bit word of y set to v, in stream order (initialisation of buffers that kernels on the same
// stream use next: a blocking memset on the null stream is not ordered against a non-blocking stream)
void launch_fill32(void* y, uint32_t v, size_t nwords, hipStream_t s);
// the rebuild predicate of a Verlet list, evaluated on the device by every build kernel: forced, or
// the largest squared site movement measured since the last build (bits in *dmax) above lim2
struct VGate {
    const unsigned long long* dmax;
    double lim2;
    int force;
    __device__ __forceinline__ bool on() const { return force || __longlong_as_double((long long)ctrl_load(dmax)) > lim2; }
};
// Verlet lists of the real-space Ewald sum, one list for all beads (built on the bead centroids with
// each site's bead spread); one cache per bead set, per-bead arrays sized for nb_max beads
#define VL_RESORT 8  // list builds per centroid sort (the others keep the last order)
#ifndef EW_SKIN_A
#define EW_SKIN_A 0.7  // skin of the real-space Ewald Verlet list (angstrom)
#endif
#define EW_BPC 6     // FP32 real-space pair blocks per CU when the k-space kernels run beside them
#define VLS 4   // lanes per site in the Verlet-list walk
#define VLG 16  // entries per list group: VLS lanes x 4 consecutive entries of each
struct VList {
    int* perm;       // [na] sorted index -> site index (of the last build)
    int* iperm;      // [na] site index -> sorted index
    int* cnt;        // [na] list length of sorted site k
    int* nex;        // [na] leading list entries that are the site's own molecule
    int* nl;         // [cap / VLG + 1][na][VLS][4] sorted neighbours j as byte offsets 16 j: entry n of k in
                     // group n / VLG, lane n % VLS, slot (n / VLS) % 4
    double* cb;      // [4][na] bead centroid and spread of each site at the last build (site order)
    double4* srt;    // [nb][na + 1] current coordinates and charge in sorted order; entry na of a bead is
                     // the charge-free pad site (zero coordinates)
    uint4* srf;      // [nb][na + 1] wrapped fractional coordinates in 32-bit fixed point (units of 2^-32,
                     // low bit clear) and the FP32 bits of the charge, sorted order; entry na of a bead is
                     // the charge-free pad site (odd coordinates, so never at distance zero)
    double* cu;      // [4][na] centroids and spreads of the build, site order
    float4* cs;      // [na] wrapped fractional centroid and spread, sorted order
    double* bb;      // [nclu][8] cluster boxes (centre, half-width, largest spread)
    int* ccnt;       // [nclu] cluster-pair list lengths
    int* clist;      // [nclu][nclu] cluster-pair lists
    int* gate;       // [2]: [1] list overflow (sticky)
    unsigned long long* dmax;  // largest squared excursion of a bead site beyond its build sphere (bits), this call
    int* nreb;       // [1] builds so far: the sites are sorted again on every VL_RESORT-th
    int cap;
};
// FP32 real-space pair functions du(s) = erfc(a r)/r and dur(s) = (sfac e^{-a^2 r^2} - du)/r^2 of
// s = r^2 in the kernel's distance units, tabulated over the scaled abscissa s' = c s, c = 2^14 / rc^2,
// so that the cut-off is s' = 2^14 and the table's PT_OCT octaves run from s' = 2 (biased FP32
// exponents 128 .. 140: their low 4 bits and the top PT_M mantissa bits are the interval, taken by a
// mask), 2^PT_M cubic intervals per octave in the remaining mantissa bits, then one interval of zeros
// at s' = 2^14 (every pair at or past the cut-off, after clamping).  The table reaches down to r
// between rc/91 and rc/64; closer pairs are evaluated without it.
#define PT_M 5
#define PT_OCT 13
#define PT_N (PT_OCT * (1 << PT_M) + 1)
static_assert(PT_OCT < 15, "the interval index holds 4 exponent bits");
struct PairTab {
    float4* d = nullptr;  // [2][PT_N] device: cubic coefficients of du (x 2^ku), then of dur (x 2^kr),
                          // in w = the low 23 - PT_M mantissa bits of s'
    float c = 0.0f;       // s' = c s
    int ku = 0, kr = 0;   // power-of-two scales of the two tables
    double key[3] = {0.0, 0.0, 0.0};  // (alpha, rc, distance unit) it was built for
};
// fill t (host part) and h with the table for Ewald parameter alpha and cut-off rc, distances in
// units of sr (the unit box: 2^-32)
void pair_table_fill(double alpha, double rc, double sr, PairTab& t, float4* h);
size_t pair_table_bytes();
// site blocks of 256 / VLS sites per bead, padded to a multiple of 8 (the XCD mapping of the real-space pair kernels)
__host__ __device__ inline int ewald_real_parts(int na) { return ((na * VLS + 255) / 256 + 7) / 8 * 8; }
int ewald_nclu(int na);
int ewald_real_slots();
int vlist_cap(int na, Box sb, double rl);
void launch_sites_vl(const double* pos, double* site, int na, int nb, double alpha, int unit, Box box,
                     const double lat[3], double* ph, const double* q, const VList& v, int gather, Box sb,
                     hipStream_t s);
// the phase factors of launch_sites_vl alone (ph = nullptr there skips them)
void launch_phases(const double* pos, int na, int nb, double alpha, int unit, Box box, const double lat[3], double* ph,
                   hipStream_t s);
void launch_vlist(const VList& v, const double* site, const double* q, int na, int nb, Box sb, CellGrid g, double rc,
                  double skin, int force, int* cell_of, int* count, int* start, hipStream_t s);
// bpc > 0: FP32 kernel capped at bpc blocks per CU (for a launch beside other streams' kernels)
// FP32 (T = float) reads the pair-function table tab, built for (alpha, rcut, sb)
template <typename T>
void launch_ewald_real(const VList& v, int na, int nb, Box sb, double rcut, double alpha, double* grad, double* parts,
                       hipStream_t s, int bpc = 0, const PairTab* tab = nullptr);
// smeared charges: adds the Gaussian-charge part of the real-space pairs to what launch_ewald_real
// wrote (gradient slot and block sums); wid = inverse widths {M-M, H-H, M-H} in the box's units
void launch_ewald_smear(const VList& v, int na, int nb, Box sb, double rcut, const double wid[3], double* grad,
                        double* parts, hipStream_t s);
int kspace_nchunk(int na);

// Replica batches: one replica's inputs of the four dominant mixed-precision force kernels (the FP32
// pair walk on the unit box with its table, k_sk_h16, k_sk_final, k_kforce_h16), as the solo
// launches take them
struct FBItem {
    const VList* vl;
    int na, nb;
    Box sb;
    double rcut, alpha;
    double *grad, *parts;  // the pair walk's gradient slot and energy/virial parts
    const PairTab* tab;
    const double *ph, *q;
    const KSet* ks;
    double* skp;
    const unsigned char* kmask;
    double *gk, *gks, *parts_k, *kpart;
};
struct ForceBatch;
ForceBatch* force_batch_new();
void force_batch_free(ForceBatch* fb);
// the launch records of n replicas, uploaded (on s, synchronously) when they differ from the last
// ones: 0 ready, 1 they differ while s is being captured (nothing uploaded), -1 HIP error
int force_batch_set(ForceBatch* fb, const FBItem* items, int n, hipStream_t s, bool capturing);
// every replica's four kernels as four launches: the pair walk on sr, the k-space chain on sk (the
// same stream: one after the other); bpc > 0 caps the pair walk at bpc blocks per CU
void force_batch_launch(const ForceBatch* fb, hipStream_t sk, hipStream_t sr, int bpc);
int kspace_ngroups(int kmax);
void launch_sk(const double* ph, const double* q, int na, int nb, const KSet& ks, double* sk_part, int fp32,
               hipStream_t s);
void launch_sk_final(int na, int nb, const KSet& ks, const unsigned char* dmask, const double* sk_part, double* gk,
                     double* gks, double* parts_k, int fp32, hipStream_t s);
int launch_kforce(const double* ph, const double* q, int na, int nb, const KSet& ks, const double* gk,
                  const double* gks, double* kpart, int fp32, hipStream_t s);
// kmax 16 .. 31: structure factor, weights and force pass in FP64 on 64 x 64 tiles (either precision
// mode); returns the partial-gradient slots written
int launch_kspace_big(const double* ph, const double* q, int na, int nb, const KSetBig& ks, const unsigned char* dmask,
                      double* skp, double* gk, double* parts_k, double* kpart, hipStream_t s);
void launch_nlist(const double* cen, int na, int njump, Box box, double cut, int* nnb, int* nbl, int* overflow,
                  hipStream_t s);
void launch_nlist_vl(const double* cen, int na, Box box, double cut, double skin, int force, int* vnnb, int* vnbl,
                     double* cb, unsigned long long* dmax, int* overflow, int* nnb, int* nbl, hipStream_t s);
// chain_parts: also apply the Ewald chain rule (launch_chain_rule) in the same launch; lj: store the
// O-O gradient per molecule there ([bead][dim][molecule]) instead of adding it to grad
template <typename T>
void launch_oo(const double* pos, int na, int nb, Box box, const int* nnb, const int* nbl, int use_list, double rcut,
               const rpmd_params& p, double* grad, double* parts, hipStream_t s, const double* chain_parts = nullptr,
               int chain_npart = 0, double chain_alpha = 1.0, double chain_scale = 1.0, int chain_acc = 0,
               unsigned long long* dmax = nullptr, double* lj = nullptr);
int oo_blocks(int na);
void launch_tem(const double* site, const double* q, int na, int nb, Box box, const int* nnb, const int* nbl,
                int use_list, double sig, double* gsite, double* parts, hipStream_t s);
int tem_blocks(int na);
void launch_intra(const double* pos, int na, int nb, int model, const rpmd_params& p, double* grad, double* parts,
                  int acc, hipStream_t s);
int intra_blocks(int na);

// ---- k_step.hip -------------------------------------------------------------------------------
void launch_nm(const double* in, double* out, const double* M, int n, size_t ne, hipStream_t s);
void launch_kick(double* p, const double* g1, double a1, const double* g2, double a2, size_t n, hipStream_t s,
                 uint64_t* ctr = nullptr);
void launch_nm_kick2(const double* p, const double* g1, double a1, const double* g2, double a2, double* pnm,
                     const double* r, double* rnm, const double* F, int n, size_t ne, hipStream_t s);
void launch_nm_com2(const double* rnm, double* rst, const double* pnm, double* p, const double* com, int nblk,
                    double inv_mtot, const double* mass, int na, const double* B, int n, size_t ne, hipStream_t s);
void launch_free_ring(double* p, double* r, const double* monod, int na, int nb, hipStream_t s,
                      const double* com = nullptr, int nblk = 0, double inv_mtot = 0.0, const double* mass = nullptr);
// launch_free_ring (COM of com_in removed from mode 0) then launch_pile over all modes, one launch;
// the kick's partial centroid sums go to com_out (not com_in)
void launch_free_ring_pile(double* p, double* r, const double* monod, int na, int nb, const double* com_in,
                           double inv_mtot, const double* mass, const double* c1, const double* c3,
                           const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add,
                           double* com_out, hipStream_t s);
int pile_blocks(int na);
// the fused PILE-L step up to the forces (kick, transform, kick, free ring, kick, inverse transform,
// centroid COM removals) for nb <= 64 in two launches; returns 0 (nothing launched) for larger nb.
// Device stream: both kicks from the draws of stream sadd1 (cosine and sine deviates).
// parts: [pile_step_blocks(na)][6]
int pile_step_blocks(int na);
int launch_pile_step(const double* p_in, const double* g1, double a1, const double* g2, double a2, const double* r_in,
                     double* p_out, double* r_out, const double* F, const double* B, int n, int na,
                     const double* monod, const double* mass, double inv_mtot, const double* c1, const double* c3,
                     const double* noise1, const double* noise2, uint64_t seed, const uint64_t* stream,
                     uint64_t sadd1, double* parts, hipStream_t s);
void launch_pile(double* p, const double* c1, const double* c3, int na, int nb, int k0, const double* noise,
                 uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* com_parts, hipStream_t s);
int bussi_parts_size();
void launch_bussi(double* p, const double* mass, int na, int nbq, long long ndof, double betan, double c,
                  const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* parts,
                  double* alpha, double* com_parts, hipStream_t s);
int gle_parts_size(int na, int nb);
void launch_gle(double* p, double* gp, const double* TS, const double* mass, int na, int nb, int n1,
                const double* noise, uint64_t seed, const uint64_t* stream, uint64_t stream_add, double* parts,
                double* pcom, double mm, hipStream_t s);
void launch_gle_draw(double* gp, double sc, int na, int nb, int n1, uint64_t seed, uint64_t stream, hipStream_t s);
void launch_com_remove(double* p, const double* mass, const double* com_parts, int nblk, double inv_mtot, int na,
                       hipStream_t s);
void launch_langevin(double* p, const double* c3m, double c1, int na, int nb, const double* noise, uint64_t seed,
                     const uint64_t* stream, uint64_t stream_add, hipStream_t s);
void launch_mode_copy(double* dst, const double* src, const int* map, int ndst, int na, double scale, hipStream_t s);
void launch_scale(double* y, double a, size_t n, hipStream_t s);
void launch_counter_inc(uint64_t* c, hipStream_t s);
void launch_sample_rpmd(double* p, const double* mass, int na, int nb, double beta, double totmass,
                        const double* noise, uint64_t seed, uint64_t stream, double* parts, double* v,
                        hipStream_t s);

// estimator sums of the state: parts[estimator_blocks(na)][10] = {kinetic, springs, centroid-virial
// sums of g1 and g2 per xyz, 0, 0} (reduce with launch_reduce_ev)
int estimator_blocks(int na);
// the barostat's move of every molecule by its centre-of-atoms sum times (scale_d - 1), SoA beads
void launch_baro_shift(double* r, int na, int nb, const double scale[3], hipStream_t s);
void launch_estimators(const double* p, const double* r, const double* g1, const double* g2, const double* mass,
                       int na, int nb, double wn2, double* parts, hipStream_t s);

// ---- k_rigid.hip ------------------------------------------------------------------------------
int rigid_blocks(int na);
// len2: the squared target lengths of the O-H1, O-H2 and H1-H2 constraints; tol: the RATTLE tolerance
// all nsub bead sub-steps of a rigid step in one launch (the gradient at entry: bead_forces of r)
void launch_rig_substeps(double* p, double* r, const double* mass, int na, int nb, int nsub, double hs, double sdt,
                         double omegan2, const double len2[3], double tol, int* err, hipStream_t s);
void launch_rig_final(double* p, const double* r, double* dvdr, const double* mass, int na, int nb, double hdt,
                      double dt, const double len2[3], double tol, int* err, double* vparts, double* vir_out,
                      hipStream_t s);
// the velocity stage of RATTLE alone (after a thermostat kick of the rigid step); scaled: p holds
// mass-scaled momenta (a GLE auxiliary channel)
void launch_rig_project(double* p, const double* r, const double* mass, int na, int nb, const double len2[3],
                        double tol, int scaled, int* err, hipStream_t s);

}  // namespace rpmd
