// capi.hip -- the C-ABI of include/rpmd_mi355x.h: context, host-side set-up, and the device
// pipelines of forces / full_forces / realft / evolve_pi (reference file:line cited per routine).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

using namespace rpmd;

#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return RPMD_ERR_HIP; \
    } while (0)
#define RETCHK(x)            \
    do {                     \
        int r_ = (x);        \
        if (r_) return r_;   \
    } while (0)

namespace {

// bead transform matrices for one n: F = R2HC/sqrt(n), B = HC2R/sqrt(n) (src/fourier.f90:42-59,
// FFTW's documented half-complex definitions), row-major [k][j]; built in long double.  Either
// pointer may be null; the outputs hold n*n doubles.
void realft_matrices_into(int n, double* F, double* B) {
    if (F)
        for (size_t i = 0; i < (size_t)n * n; i++) F[i] = 0.0;
    if (B)
        for (size_t i = 0; i < (size_t)n * n; i++) B[i] = 0.0;
    const long double pi = acosl(-1.0L);
    const long double sc = sqrtl(1.0L / n);
    for (int j = 0; j < n && F; j++) {
        for (int k = 0; k <= n / 2; k++) {
            long double a = 2.0L * pi * (long double)(((long long)j * k) % n) / n;
            F[(size_t)k * n + j] = (double)(sc * cosl(a));
        }
        for (int k = 1; k <= (n - 1) / 2; k++) {
            long double a = 2.0L * pi * (long double)(((long long)j * k) % n) / n;
            F[(size_t)(n - k) * n + j] = (double)(-sc * sinl(a));
        }
    }
    for (int j = 0; j < n && B; j++) {
        B[(size_t)j * n + 0] = (double)sc;
        for (int k = 1; k <= (n - 1) / 2; k++) {
            long double a = 2.0L * pi * (long double)(((long long)j * k) % n) / n;
            B[(size_t)j * n + k] = (double)(2.0L * sc * cosl(a));
            B[(size_t)j * n + (n - k)] = (double)(-2.0L * sc * sinl(a));
        }
        if (n % 2 == 0) B[(size_t)j * n + n / 2] = (double)((j % 2 == 0 ? 1.0L : -1.0L) * sc);
    }
}
void realft_matrices(int n, std::vector<double>& F, std::vector<double>& B) {
    F.assign((size_t)n * n, 0.0);
    B.assign((size_t)n * n, 0.0);
    realft_matrices_into(n, F.data(), B.data());
}

struct NMMat {
    double *F = nullptr, *B = nullptr;    // plain
    double *Fs = nullptr, *Bs = nullptr;  // contraction-scaled: B*sqrt(n/nb), F*sqrt(nb/n)
    int scaled_nb = -1;
    // contraction mode maps for (n, scaled_nb), src/evolve_pi.f90:384-401,417-425: mi[n] (kept mode
    // of each contracted mode), then mo[scaled_nb] (contracted mode of each full mode, -1: dropped)
    int* maps = nullptr;
};

// the buffers whose roles a step exchanges (r <-> rst every step; dvdr or dvdrl <-> tmp in
// contracted steps): a captured graph replays correctly only with the roles it was captured with
struct Roles {
    double* a[5];
    bool operator==(const Roles& o) const { return !memcmp(a, o.a, sizeof a); }
    bool operator!=(const Roles& o) const { return !(*this == o); }
};

}  // namespace

struct rpmd_ctx {
    rpmd_params prm;
    int na = 0, nb_max = 0, prec = 0;
    hipStream_t st = nullptr;
    // force pipeline
    double *aos = nullptr, *pos = nullptr, *grad = nullptr, *site = nullptr, *gsite = nullptr;
    double *q = nullptr, *mass = nullptr, *cen = nullptr, *parts = nullptr, *ev = nullptr, *skp = nullptr, *gk = nullptr, *gks = nullptr;
    double *ph = nullptr, *fpart = nullptr;
    // real-space Verlet-list caches: two bead sets (e.g. all beads and RP-contracted ones) for the
    // device-resident state [0, 1] and two for host-synchronous calls [2, 3], so that a forces call
    // never moves the rebuilds (and with them the summation order) of a state's lists
    struct VCache {
        VList v{};
        bool alloc = false;
        int nb = -1;
        double key[5] = {0, 0, 0, 0, 0};  // box, rc, skin
        long long last_use = 0;
    } vc[4];
    long long vc_clock = 0;
    int *cell_of = nullptr, *cnt = nullptr, *start = nullptr, *perm = nullptr, *nnb = nullptr, *nbl = nullptr,
        *ovf = nullptr;
    // Verlet O-O centroid list of the device-resident calls
    int *oo_nnb = nullptr, *oo_nbl = nullptr, *oo_ovf = nullptr;
    double* oo_cb = nullptr;
    unsigned long long* oo_dmax = nullptr;
    double oo_key[5] = {0, 0, 0, 0, 0};
    bool oo_built = false;
    unsigned char* kmask = nullptr;
    size_t off_real = 0, off_k = 0, off_oo = 0, off_tem = 0, off_intra = 0, nparts = 0;
    int ncell_cap = 0;
    // step state
    double *p = nullptr, *r = nullptr, *dvdr = nullptr, *dvdr2 = nullptr, *dvdrl = nullptr, *dvdre = nullptr,
           *rst = nullptr, *pnm = nullptr, *rnm = nullptr, *tmp = nullptr, *rb = nullptr, *dvdrb = nullptr,
           *tmp2 = nullptr;
    double *monod = nullptr, *pile_c1 = nullptr, *pile_c3 = nullptr, *com_parts = nullptr, *lang_c3m = nullptr;
    double* com_parts2 = nullptr;  // a second set of centroid partial sums (read and written in one launch)
    double* pstep_parts = nullptr;  // [pile_step_blocks(na)][6] centroid sums of the fused PILE step
    double* noise = nullptr;
    size_t noise_cap = 0;
    uint64_t* step_ctr = nullptr;
    uint64_t seed = 0;
    // thermostat state: Bussi partial sums and scale; GLE auxiliary momenta gp (n1 channels of
    // 3 na nb, channel 0 = mass-scaled p), propagators T and S, partial sums
    double *bussi_parts = nullptr, *bussi_alpha = nullptr;
    double *gle_gp = nullptr, *gle_TS = nullptr, *gle_parts = nullptr, *gle_pcom = nullptr;
    int gle_n1 = 0, gle_nb = 0;
    bool lf_merged = false;  // the last step's Ewald and O-O forces came from one forces(iopt = 1) call
    // rigid molecules: RATTLE failure flag, virial partials
    int* rat_err = nullptr;
    double* rig_vparts = nullptr;
    std::map<int, NMMat> nm;
    // host staging (pinned)
    double* h_stage = nullptr;
    size_t h_cap = 0;
    // FP32 real-space pair-function tables: [0] physical box, [1] unit box (host copies kept for the
    // upload)
    PairTab ptab[2];
    std::vector<float4> ptab_h[2];
    // cached k set
    KSet ks{};
    double ks_key[8] = {0};
    bool ks_valid = false;
    // kmax 16 .. 31: the k set on 64 x 64 tiles and the FP64 buffers of its kernels (allocated on
    // first use): mask, chunk partials, G tiles, energy/virial parts, gradient slots
    KSetBig ksb{};
    double ksb_key[8] = {0};
    bool ksb_valid = false;
    unsigned char* kmask_big = nullptr;
    double *skp_big = nullptr, *gk_big = nullptr, *parts_big = nullptr, *fpart_big = nullptr;
    size_t skp_big_n = 0, gk_big_n = 0;
    double sumq2 = 0.0, mtot = 0.0;
    std::vector<double> hq;
    // device-resident state
    int st_nb = 0;
    Box st_box{};
    double st_beta = 0, st_dt = 0;
    bool st_valid = false;
    std::vector<double> st_z;  // the state's charges (host calls may upload others)
    std::vector<double> hmass;
    // energies of the last step: slots [HF, EW, LJ, TEM, EWC] x 5 terms x 10
    double* ev_step = nullptr;
    double vir_hf[9] = {0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int use_graph = 0;
    // epoch: advanced by every call that may change what a captured step reads (tables, caches,
    // charges, parameters, the state); a graph (this context's or a replica batch's) records the
    // epoch and the buffer roles at capture and is replayed only while both still hold
    uint64_t epoch = 1, g_epoch = 0;
    Roles g_roles{};
    // second stream for the real-space Ewald branch of device-resident force calls (streams >= 2):
    // forked after the sites, joined before the chain rule; third stream (streams = 3) for the O-O
    // list and pair sums of those calls and the intramolecular forces of a step, joined before the
    // chain rule; default priority, as the first
    hipStream_t st2 = nullptr, st3 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork3 = nullptr, ev_join3 = nullptr;
    int streams = 2;
    double* goo = nullptr;  // O-O gradient of the oxygens per molecule, [bead][dim][molecule]
    // barostat (module barostat, src/module_barostat.f90:1-5): RPMD_BARO_*, taub, patm, dv and the
    // MC counters; the library's own stream of the MC draws; device sums of virial_ke per substep
    int baro = 0;
    double taub = 0.0, patm = 0.0, dv = 0.0;
    long long nacc = 0, nmove = 0;
    uint64_t baro_draws = 0;
    double *baro_parts = nullptr, *baro_ev = nullptr;
    int baro_mts = 0;
    // kernel timing with HIP events on the context stream (eager launches only)
    int prof = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, size_t>> ev_marks;  // (kernel id, index of the start event)
};

namespace {

Roles roles_of(const rpmd_ctx* c) { return Roles{{c->r, c->rst, c->dvdr, c->dvdrl, c->tmp}}; }

void drop_graph(rpmd_ctx* c) {
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// every captured graph that reads this context (its own, a replica batch's) is stale from here on
void invalidate(rpmd_ctx* c) {
    c->epoch++;
    drop_graph(c);
}

// timed launch bracket: kernel ids 0 ewald_real, 1 k_sk, 2 k_kforce, 3 k_sk_final
struct ProfScope {
    rpmd_ctx* c;
    int id;
    size_t i0;
    ProfScope(rpmd_ctx* c_, int id_) : c(c_), id(id_), i0(0) {
        if (!c->prof) return;
        i0 = c->ev_marks.size() * 2;
        while (c->ev_pool.size() < i0 + 2) {
            hipEvent_t e;
            hipEventCreate(&e);
            c->ev_pool.push_back(e);
        }
        hipEventRecord(c->ev_pool[i0], c->st);
    }
    ~ProfScope() {
        if (!c->prof) return;
        hipEventRecord(c->ev_pool[i0 + 1], c->st);
        c->ev_marks.push_back({id, i0});
    }
};

Box make_box(const double b[3]) {
    Box x;
    for (int d = 0; d < 3; d++) {
        x.L[d] = b[d];
        x.inv[d] = 1.0 / b[d];
    }
    return x;
}
Box unit_box() {
    Box x;
    for (int d = 0; d < 3; d++) x.L[d] = x.inv[d] = 1.0;
    return x;
}

template <typename T>
int dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T)) == hipSuccess ? 0 : RPMD_ERR_NOMEM;
}

int intra_model(const rpmd_params& P) {
    if (P.intra_model != RPMD_INTRA_AUTO) return P.intra_model;
    if (P.alp == 0.0) return RPMD_INTRA_HARMONIC;  // src/md_forces.f90:254-256
    return P.alpb != 0.0 ? RPMD_INTRA_MORSE : RPMD_INTRA_MORSE_HARM;
}

// the reference's k set as a mask [kx][ky+15][kz+15] (src/ewald_cubic.f90:292-520 loop exits;
// src/ewald_noncubic.f90:385-397 strict comparisons)
// the reference's k set (src/ewald_cubic.f90:235-533: cubic, the loops with their breaks;
// src/ewald_noncubic.f90:314-457: every (kx, ky, kz) with 0 < k^2 < rkmax^2, kx >= 0) on tiles of
// T x T entries per kx (index k + T / 2), kx in [0, H]: the mask m and the per-row bounds of KS
template <typename KS>
void kset_fill(const rpmd_params& P, const double box[3], int T, int H, KS& ks, std::vector<unsigned char>& m) {
    const int K = P.kmax, O = T / 2;
    m.assign((size_t)(H + 1) * T * T, 0);
    auto M = [&](int kx, int ky, int kz) -> unsigned char& { return m[((size_t)kx * T + ky + O) * T + kz + O]; };
    const double pi = acos(-1.0);
    ks.kmax = K;
    const double rkmax2 = P.rkmax * P.rkmax;
    if (P.iamcub) {
        const double rlat = 2.0 * pi;
        for (int kx = 0; kx <= K; kx++) {
            double rkx = rlat * kx, rky, rkz, rk2 = rkx * rkx;
            if (rk2 > 0.0) M(kx, 0, 0) = 1;
            for (int kz = 1; kz <= K; kz++) {
                rkz = rlat * kz;
                rk2 = rkx * rkx + rkz * rkz;
                if (rk2 > rkmax2) break;
                M(kx, 0, kz) = M(kx, 0, -kz) = 1;
            }
            for (int ky = 1; ky <= K; ky++) {
                rky = rlat * ky;
                rk2 = rkx * rkx + rky * rky;
                if (rk2 > rkmax2) break;
                M(kx, ky, 0) = M(kx, -ky, 0) = 1;
                for (int kz = 1; kz <= K; kz++) {
                    rkz = rlat * kz;
                    rk2 = rkx * rkx + rky * rky + rkz * rkz;
                    if (rk2 > rkmax2) break;
                    M(kx, ky, kz) = M(kx, ky, -kz) = M(kx, -ky, kz) = M(kx, -ky, -kz) = 1;
                }
            }
        }
        ks.lat[0] = ks.lat[1] = ks.lat[2] = rlat;
        ks.f0 = rlat;
    } else {
        const double twopi = 2.0 * pi;
        const double xl = twopi / box[0], yl = twopi / box[1], zl = twopi / box[2];
        for (int kx = 0; kx <= K; kx++)
            for (int ky = -K; ky <= K; ky++)
                for (int kz = -K; kz <= K; kz++) {
                    double rkx = xl * kx, rky = yl * ky, rkz = zl * kz;
                    double rk2 = rkx * rkx + rky * rky + rkz * rkz;
                    if (rk2 < rkmax2 && rk2 != 0.0) M(kx, ky, kz) = 1;
                }
        ks.lat[0] = xl;
        ks.lat[1] = yl;
        ks.lat[2] = zl;
        ks.f0 = twopi / (box[0] * box[1] * box[2]);
    }
    ks.b = 0.25 / (P.walpha * P.walpha);
    ks.cubic = P.iamcub ? 1 : 0;
    for (int kx = 0; kx <= H; kx++) {
        int kyb = -1, kzb = 0;
        for (int ky = -(O - 1); ky <= O - 1; ky++)
            for (int kz = -(O - 1); kz <= O - 1; kz++)
                if (kx <= K && M(kx, ky, kz)) {
                    kyb = std::max(kyb, std::abs(ky));
                    kzb = std::max(kzb, std::abs(kz));
                }
        ks.ky_max[kx] = kyb;
        ks.kz_bound[kx] = kzb;
        for (int iy = 0; iy < T; iy++) {
            int r = -1;
            for (int iz = 0; iz < T; iz++)
                if (kx <= K && m[((size_t)kx * T + iy) * T + iz]) r = std::max(r, std::abs(iz - O));
            ks.kz_row[kx][iy] = (signed char)r;
        }
    }
}

int build_kset(rpmd_ctx* c, const double box[3]) {
    const rpmd_params& P = c->prm;
    double key[8] = {box[0], box[1], box[2], P.walpha, P.rkmax, (double)P.kmax, (double)P.iamcub, 0};
    if (c->ks_valid && !memcmp(key, c->ks_key, sizeof key)) return 0;
    if (P.kmax > RPMD_KMAX_DEV || P.kmax < 0) return RPMD_ERR_KMAX;
    std::vector<unsigned char> m;
    KSet ks{};
    kset_fill(P, box, RPMD_KTILE, RPMD_KMAX_DEV, ks, m);
    HIPCHK(hipMemcpyAsync(c->kmask, m.data(), m.size(), hipMemcpyHostToDevice, c->st));
    c->ks = ks;
    memcpy(c->ks_key, key, sizeof key);
    c->ks_valid = true;
    return 0;
}

// kmax 16 .. 31: the k set on 64 x 64 tiles, its mask on the device and the big-tile buffers (first
// use, or a larger kmax, allocates them: always in an eager call, the step before a capture)
int build_kset_big(rpmd_ctx* c, const double box[3]) {
    const rpmd_params& P = c->prm;
    double key[8] = {box[0], box[1], box[2], P.walpha, P.rkmax, (double)P.kmax, (double)P.iamcub, 0};
    if (c->ksb_valid && !memcmp(key, c->ksb_key, sizeof key)) return 0;
    if (P.kmax > RPMD_KMAX_BIG || P.kmax <= RPMD_KMAX_DEV) return RPMD_ERR_KMAX;
    const int na = c->na, nbl = c->nb_max, T = RPMD_KTILE_BIG;
    const size_t skp_n = (size_t)nbl * (P.kmax + 1) * kspace_nchunk(na) * T * T * 2;
    const size_t gk_n = (size_t)nbl * (P.kmax + 1) * T * T * 2;
    for (hipStream_t q : {c->st, c->st2, c->st3}) HIPCHK(hipStreamSynchronize(q));
    if (c->skp_big_n < skp_n) {
        if (c->skp_big) hipFree(c->skp_big);
        c->skp_big = nullptr;
        c->skp_big_n = 0;
        RETCHK(dalloc(&c->skp_big, skp_n) ? RPMD_ERR_NOMEM : 0);
        c->skp_big_n = skp_n;
    }
    if (c->gk_big_n < gk_n) {
        if (c->gk_big) hipFree(c->gk_big);
        c->gk_big = nullptr;
        c->gk_big_n = 0;
        RETCHK(dalloc(&c->gk_big, gk_n) ? RPMD_ERR_NOMEM : 0);
        c->gk_big_n = gk_n;
    }
    if (!c->kmask_big) HIPCHK(hipMalloc((void**)&c->kmask_big, (size_t)(RPMD_KMAX_BIG + 1) * T * T));
    if (!c->parts_big) RETCHK(dalloc(&c->parts_big, (size_t)nbl * (RPMD_KMAX_BIG + 1) * 10) ? RPMD_ERR_NOMEM : 0);
    if (!c->fpart_big)
        RETCHK(dalloc(&c->fpart_big,
                      (size_t)(ewald_real_slots() + kspace_ngroups(RPMD_KMAX_BIG)) * 3 * na * nbl) ? RPMD_ERR_NOMEM
                                                                                                  : 0);
    std::vector<unsigned char> m;
    KSetBig ks{};
    kset_fill(P, box, T, RPMD_KMAX_BIG, ks, m);
    HIPCHK(hipMemcpyAsync(c->kmask_big, m.data(), m.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->ksb = ks;
    memcpy(c->ksb_key, key, sizeof key);
    c->ksb_valid = true;
    invalidate(c);
    return 0;
}

// the FP32 pair-function table for (alpha, rc) in the units of box sb, built and uploaded when the
// parameters change (never under stream capture: a captured step follows an eager one with the same
// parameters).  The context's streams are drained first, as queued kernels may read the old table.
static int pair_table_for(rpmd_ctx* c, const Box& sb, double rc, double alpha, hipStream_t s, const PairTab** out) {
    const int unit = sb.L[0] == 1.0 && sb.L[1] == 1.0 && sb.L[2] == 1.0;
    const double sr = unit ? 0x1p-32 : 1.0;
    PairTab& t = c->ptab[unit];
    *out = &t;
    if (t.d && t.key[0] == alpha && t.key[1] == rc && t.key[2] == sr) return 0;
    const size_t bytes = pair_table_bytes();
    if (!t.d) HIPCHK(hipMalloc((void**)&t.d, bytes));
    for (hipStream_t q : {s, c->st, c->st2, c->st3}) HIPCHK(hipStreamSynchronize(q));
    std::vector<float4>& h = c->ptab_h[unit];
    h.resize(bytes / sizeof(float4));
    pair_table_fill(alpha, rc, sr, t, h.data());
    HIPCHK(hipMemcpyAsync(t.d, h.data(), bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

// Morton (power-of-two) bin count of a cell grid
int morton_bins(const CellGrid& g) {
    int p = 1;
    while (p < g.nc[0] || p < g.nc[1] || p < g.nc[2]) p <<= 1;
    return p * p * p;
}

// Morton cell grid of the cluster sort: about 4 sites per cell, cells near-cubic in space
CellGrid make_grid(const Box& sb, int na) {
    CellGrid g;
    const double h = cbrt(sb.L[0] * sb.L[1] * sb.L[2] * 4.0 / std::max(na, 1));
    for (int d = 0; d < 3; d++) g.nc[d] = std::max(1, std::min(64, (int)lround(sb.L[d] / h)));
    g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
    return g;
}

// the bead-transform matrices for n beads (allocated on first use, outside stream capture); with
// nb_full > 0 also the contraction-scaled pair and the mode maps for nb_full physical beads
NMMat* get_nm(rpmd_ctx* c, int n, int nb_full) {
    auto it = c->nm.find(n);
    if (it == c->nm.end()) {
        std::vector<double> F, B;
        realft_matrices(n, F, B);
        NMMat m;
        if (dalloc(&m.F, F.size()) || dalloc(&m.B, B.size()) || dalloc(&m.Fs, F.size()) || dalloc(&m.Bs, B.size()) ||
            dalloc(&m.maps, (size_t)n + 256))
            return nullptr;
        // on the context stream (a null-stream copy is not ordered against it), complete before the
        // host vectors go
        if (hipMemcpyAsync(m.F, F.data(), F.size() * 8, hipMemcpyHostToDevice, c->st) != hipSuccess ||
            hipMemcpyAsync(m.B, B.data(), B.size() * 8, hipMemcpyHostToDevice, c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess)
            return nullptr;
        it = c->nm.emplace(n, m).first;
    }
    NMMat& m = it->second;
    if (nb_full > 0 && m.scaled_nb != nb_full) {
        std::vector<double> F, B;
        realft_matrices(n, F, B);
        const double sB = sqrt((double)n / (double)nb_full), sF = sqrt((double)nb_full / (double)n);
        for (auto& x : B) x *= sB;
        for (auto& x : F) x *= sF;
        // ring_contract keeps modes 0..h and their mirrors nb-h..nb-1 (src/evolve_pi.f90:417-425);
        // force_contract puts contracted mode k back at k and n - k at nb - k (:388-401)
        std::vector<int> mp((size_t)n + nb_full, -1);
        int* mi = mp.data();
        int* mo = mp.data() + n;
        const int h = (n - 1) / 2;
        mi[0] = 0;
        mo[0] = 0;
        for (int k = 1; k <= h; k++) {
            mi[k] = k;
            mi[n - k] = nb_full - k;
            mo[k] = k;
            mo[nb_full - k] = n - k;
        }
        if (hipMemcpyAsync(m.Fs, F.data(), F.size() * 8, hipMemcpyHostToDevice, c->st) != hipSuccess ||
            hipMemcpyAsync(m.Bs, B.data(), B.size() * 8, hipMemcpyHostToDevice, c->st) != hipSuccess ||
            hipMemcpyAsync(m.maps, mp.data(), mp.size() * sizeof(int), hipMemcpyHostToDevice, c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess)
            return nullptr;
        m.scaled_nb = nb_full;
    }
    return &m;
}

// ---- forces on device positions (SoA, nb beads); grad overwritten; ev_slot[5][10] ---------------
// ev terms: 0 Ewald real space, 1 Ewald k space (self term added on the host), 2 O-O, 3 intra,
// 4 TEM; each is the bead average.  check_lists: read back the neighbour-list status (host sync).
void free_vlist(VList& v) {
    void* bufs[] = {v.perm, v.iperm, v.cnt, v.nex, v.nl, v.cb, v.srt, v.srf, v.cu, v.cs, v.bb, v.ccnt, v.clist,
                    v.gate, v.dmax, v.nreb};
    for (void* b : bufs)
        if (b) hipFree(b);
    v = VList{};
}

void free_vcache(rpmd_ctx* c) {
    for (auto& x : c->vc) {
        if (x.alloc) free_vlist(x.v);
        x.alloc = false;
        x.nb = -1;
    }
}

// the Verlet-list cache for this bead count (least recently used slot otherwise); allocated, or
// grown, outside of stream capture only; *force = 1 when the cached lists belong to another set-up
int vlist_for(rpmd_ctx* c, int nb, const double box[3], double rc, double skin, const Box& sb, VList** out,
              int* force, int host) {
    const int na = c->na, nbl = c->nb_max, s0 = host ? 2 : 0;
    int slot = -1;
    for (int i = s0; i < s0 + 2; i++)
        if (c->vc[i].alloc && c->vc[i].nb == nb) slot = i;
    if (slot < 0) slot = c->vc[s0].last_use <= c->vc[s0 + 1].last_use ? s0 : s0 + 1;
    auto& x = c->vc[slot];
    x.last_use = ++c->vc_clock;
    const double key[5] = {box[0], box[1], box[2], rc, skin};
    *force = (x.nb != nb || memcmp(key, x.key, sizeof key) != 0) ? 1 : 0;
    // capacity for the pairs within rc + skin plus twice a bead spread of up to two skins
    const int cap = vlist_cap(na, sb, rc + 5.0 * skin);
    if (!x.alloc || x.v.cap < cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(c->st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return RPMD_ERR_ARG;
        HIPCHK(hipStreamSynchronize(c->st));
        if (x.alloc) free_vlist(x.v);
        x.alloc = false;
        VList v{};
        const int nclu = ewald_nclu(na);
        const size_t ns = (size_t)na * nbl;
        int e = 0;
        e |= dalloc(&v.perm, na);
        e |= dalloc(&v.iperm, na);
        e |= dalloc(&v.cnt, na);
        e |= dalloc(&v.nex, na);
        e |= dalloc(&v.nl, (size_t)na * (cap + VLG));  // + the padding group
        e |= dalloc(&v.cb, (size_t)4 * na);
        e |= dalloc(&v.srt, ns + nbl);  // + the pad site of each bead
        e |= dalloc(&v.srf, ns + nbl);
        e |= dalloc(&v.cu, (size_t)4 * na);
        e |= dalloc(&v.cs, na);
        e |= dalloc(&v.bb, (size_t)8 * nclu);
        e |= dalloc(&v.ccnt, nclu);
        e |= dalloc(&v.clist, (size_t)nclu * nclu);
        e |= dalloc(&v.gate, 2);
        e |= dalloc(&v.dmax, 1);
        e |= dalloc(&v.nreb, 1);
        if (e) {
            free_vlist(v);
            return RPMD_ERR_NOMEM;
        }
        v.cap = cap;
        // build centroids NaN (all bits set): the first call always builds.  Filled on the
        // context stream, ahead of the kernels that use them (a null-stream memset is not ordered
        // against this non-blocking stream and could land after the first build)
        launch_fill32(v.cb, 0xffffffffu, (size_t)2 * 4 * na, c->st);
        launch_fill32(v.gate, 0u, 2, c->st);
        launch_fill32(v.dmax, 0u, 2, c->st);
        launch_fill32(v.nreb, 0u, 1, c->st);
        x.v = v;
        x.alloc = true;
        *force = 1;
    }
    x.nb = nb;
    memcpy(x.key, key, sizeof key);
    *out = &x.v;
    return 0;
}

// defer: append the energy reductions (outputs relative to c->ev_step) instead of launching them
// main_st: run the call on this stream instead of the context stream (unforked calls only).
// intra_grad (with defer, iopt 1): also the intramolecular forces of pos into intra_grad, with their
// energy in component 3 of the slot at c->ev_step (as a separate iopt 4 call there would); returns
// 1 through *intra_done when it did so (on the second stream, after the O-O terms), else 0
int forces_dev(rpmd_ctx* c, const double* pos, int nb, const double box[3], int iopt, double* grad, double* ev_slot,
               bool check_lists, ReduceJobs* defer = nullptr, hipStream_t main_st = nullptr,
               double* intra_grad = nullptr, int* intra_done = nullptr, KickArgs* kick = nullptr,
               int* kick_done = nullptr) {
    const rpmd_params& P = c->prm;
    const int na = c->na;
    hipStream_t s = main_st ? main_st : c->st;
    if (iopt < 0 || iopt > 5) return RPMD_ERR_IOPT;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    const Box bx = make_box(box);
    const double boxmin = std::min(box[0], std::min(box[1], box[2]));
    const double boxmax = std::max(box[0], std::max(box[1], box[2]));
    const size_t N = (size_t)3 * na * nb;
    const bool mixed = c->prec == RPMD_PREC_MIXED;
    const bool need_oo = (iopt == 3 || iopt <= 1);
    // the first kernel writing grad stores, the others add; the O-O term alone writes only the
    // oxygens, so its call starts from zero.  ev_slot components not computed here are left as
    // they are: combine_ev reads only the computed ones.
    if (iopt == 3) launch_zero(grad, N, s);
    int gacc = iopt == 3 ? 1 : 0;
    ReduceJobs jobs;
    int use_list = 0;
    unsigned long long* oo_dmax = nullptr;
    const double oo_cut = std::min(P.rcut, 0.5 * boxmin);
    // device-resident calls: a Verlet list of the O-O centroids within cut + 1 A, rebuilt when an
    // oxygen centroid has moved by about 0.5 A, from which the pairs within cut are taken
    auto oo_vlist = [&](hipStream_t so) {
        const double skin = 1.0 / 0.5291772108;
        const double key[5] = {box[0], box[1], box[2], oo_cut, skin};
        const int force = (!c->oo_built || memcmp(key, c->oo_key, sizeof key) != 0) ? 1 : 0;
        memcpy(c->oo_key, key, sizeof key);
        c->oo_built = true;
        launch_centroid_vl(pos, c->cen, na, nb, c->oo_cb, c->oo_dmax, so);
        launch_nlist_vl(c->cen, na, bx, oo_cut, skin, force, c->oo_nnb, c->oo_nbl, c->oo_cb, c->oo_dmax, c->oo_ovf,
                        c->nnb, c->nbl, so);
        use_list = 1;
        oo_dmax = c->oo_dmax;  // cleared by the O-O kernel
    };
    // device-resident calls with the Ewald sum: the O-O list and pair sums run beside the k-space
    // kernels, after the real-space pairs on the second stream (or from here on the third), their
    // gradient per molecule (goo) added by the chain-rule kernel after the join
    const bool oo_branch = need_oo && iopt <= 2 && c->streams > 1 && !check_lists && !c->prof && !main_st;
    const bool fork3 = oo_branch && c->streams > 2;
    auto oo_work = [&](hipStream_t so) {
        if (nb > 3) oo_vlist(so);
        double* po = c->parts + c->off_oo;
        if (mixed)
            launch_oo<float>(pos, na, nb, bx, c->nnb, c->nbl, use_list, P.rcut, P, grad, po, so, nullptr, 0, 1.0, 1.0,
                             0, oo_dmax, c->goo);
        else
            launch_oo<double>(pos, na, nb, bx, c->nnb, c->nbl, use_list, P.rcut, P, grad, po, so, nullptr, 0, 1.0,
                              1.0, 0, oo_dmax, c->goo);
    };
    if (fork3) {
        HIPCHK(hipEventRecord(c->ev_fork3, s));
        HIPCHK(hipStreamWaitEvent(c->st3, c->ev_fork3, 0));
        oo_work(c->st3);
        HIPCHK(hipEventRecord(c->ev_join3, c->st3));
    }
    if (need_oo && nb > 3 && !oo_branch) {
        // centroid O-O list (src/md_forces.f90:34-53)
        if (check_lists) {
            // the list of this call, and lj_driver's fallback when it is empty (src/ljones.f90:14-26)
            use_list = 1;
            launch_centroid(pos, c->cen, na, nb, s);
            launch_zero_int(c->ovf, 2, s);
            launch_nlist(c->cen, na, 3, bx, oo_cut, c->nnb, c->nbl, c->ovf, s);
            int h[2];
            HIPCHK(hipMemcpyAsync(h, c->ovf, sizeof h, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (h[0] > 0) return RPMD_ERR_NEIGH_OVERFLOW;
            if (h[1] == 0) use_list = 0;
        } else {
            oo_vlist(s);
        }
    }
    int chain_npart = 0, chain_acc = 0;
    double chain_scale = 1.0;
    double* chain_fp = c->fpart;  // the Ewald gradient slots the chain rule sums
    bool fork = false;
    if (iopt <= 2) {
        // kmax 16 .. 31: the FP64 big-tile k-space kernels (in either precision mode), with their own
        // gradient slots
        const bool bigk = P.kmax > RPMD_KMAX_DEV;
        if (bigk)
            RETCHK(build_kset_big(c, box));
        else
            RETCHK(build_kset(c, box));
        const double* lat = bigk ? c->ksb.lat : c->ks.lat;
        double* fp = bigk ? c->fpart_big : c->fpart;
        chain_fp = fp;
        const int cubic = P.iamcub;
        const Box sb = cubic ? unit_box() : bx;
        const PairTab* ptab = nullptr;
        if (mixed) RETCHK(pair_table_for(c, sb, P.wrcut, P.walpha, s, &ptab));
        CellGrid g = make_grid(sb, na);
        if (morton_bins(g) > c->ncell_cap) return RPMD_ERR_ARG;
        // list skin EW_SKIN_A (src/main.f90 uses none: the reference rebuilds its cells every call)
        const double skin = (cubic ? 1.0 / box[0] : 1.0) * (EW_SKIN_A / 0.5291772108);
        VList* vl = nullptr;
        int force = 0;
        RETCHK(vlist_for(c, nb, box, P.wrcut, skin, sb, &vl, &force, check_lists ? 1 : 0));
        // the real-space branch (sites, list upkeep and pairs) runs on the second stream, beside the
        // k-space kernels on the first, which then form the phases themselves, unless the lists are
        // checked on the host or kernels are being timed
        fork = c->streams > 1 && !check_lists && !c->prof;
        hipStream_t sr = s;
        if (fork) {
            HIPCHK(hipEventRecord(c->ev_fork, s));
            HIPCHK(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
            sr = c->st2;
            launch_phases(pos, na, nb, P.alpha, cubic, bx, lat, c->ph, s);
        }
        // sites, (phases,) list displacement and (for a built list) the sorted coordinates
        launch_sites_vl(pos, c->site, na, nb, P.alpha, cubic, bx, lat, fork ? nullptr : c->ph, c->q, *vl, !force,
                        sb, sr);
        {
            ProfScope ps(c, 4);
            launch_vlist(*vl, c->site, c->q, na, nb, sb, g, P.wrcut, skin, force, c->cell_of, c->cnt, c->start, sr);
        }
        if (check_lists) {
            int h[2];
            HIPCHK(hipMemcpyAsync(h, vl->gate, sizeof h, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (h[1]) return RPMD_ERR_NEIGH_OVERFLOW;
        }
        double* pr = c->parts + c->off_real;
        {
            ProfScope ps(c, 0);
            if (mixed)
                launch_ewald_real<float>(*vl, na, nb, sb, P.wrcut, P.walpha, fp, pr, sr, fork ? EW_BPC : 0, ptab);
            else
                launch_ewald_real<double>(*vl, na, nb, sb, P.wrcut, P.walpha, fp, pr, sr);
        }
        if (P.wm > 0.0 || P.wh > 0.0) {
            // Gaussian charges: inverse widths of the M-M, H-H and M-H pairs, in unit-box lengths for
            // the cubic form (src/ewald_cubic.f90:40-54; src/ewald_noncubic.f90:18-33)
            const double ob = cubic ? 1.0 / box[0] : 1.0;
            const double sm = sqrt(1.0 / (2.0 * P.wm)) * ob, sh = sqrt(1.0 / (2.0 * P.wh)) * ob;
            const double wid[3] = {1.0 / (2.0 * sm), 1.0 / (2.0 * sh), 1.0 / sqrt(2.0 * (sm * sm + sh * sh))};
            launch_ewald_smear(*vl, na, nb, sb, P.wrcut, wid, fp, pr, sr);
        }
        if (oo_branch && !fork3) {
            oo_work(c->st2);
            if (intra_grad && defer && intra_done) {
                double* pi = c->parts + c->off_intra;
                launch_intra(pos, na, nb, intra_model(P), P, intra_grad, pi, 0, c->st2);
                if (defer->n >= RED_JOBS) return RPMD_ERR_ARG;
                defer->add(pi, nb * intra_blocks(na), 30, 1.0 / nb);
                *intra_done = 1;
            }
        }
        if (fork) HIPCHK(hipEventRecord(c->ev_join, c->st2));
        double* pk = bigk ? c->parts_big : c->parts + c->off_k;
        int nkslots = 0;
        if (bigk) {
            nkslots = launch_kspace_big(c->ph, c->q, na, nb, c->ksb, c->kmask_big, c->skp_big, c->gk_big, pk,
                                        fp + (size_t)ewald_real_slots() * 3 * na * nb, s);
        } else {
            {
                ProfScope ps(c, 1);
                launch_sk(c->ph, c->q, na, nb, c->ks, c->skp, mixed, s);
            }
            {
                ProfScope ps(c, 3);
                launch_sk_final(na, nb, c->ks, c->kmask, c->skp, c->gk, c->gks, pk, mixed, s);
            }
            {
                ProfScope ps(c, 2);
                nkslots = launch_kforce(c->ph, c->q, na, nb, c->ks, c->gk, c->gks,
                                        c->fpart + (size_t)ewald_real_slots() * 3 * na * nb, mixed, s);
            }
        }
        if (fork) HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
        const double sc = cubic ? 1.0 / box[0] : 1.0;
        // real-space slot then the k-space slots, summed in that order; with the O-O term in this call,
        // its kernel applies the chain rule too
        chain_npart = ewald_real_slots() + nkslots;
        chain_scale = sc * sc;
        chain_acc = gacc;
        // energies: real and k parts into the Ewald slot (two reductions summed on the host)
        jobs.add(pr, nb * ewald_real_parts(na), 0, sc / nb);
        jobs.add(pk, nb * (P.kmax + 1), 10, sc / nb);
        if (oo_branch) {
            if (fork3) HIPCHK(hipStreamWaitEvent(s, c->ev_join3, 0));
            jobs.add(c->parts + c->off_oo, nb * oo_blocks(na), 20, 1.0 / nb);
            // with kick (iopt 1: grad is final here), the step's closing kick in the same launch, and
            // with defer the step's energy reductions (every part is written by now) as extra blocks
            const bool fuse = kick && kick_done && iopt == 1;
            if (fuse && defer) {
                const int base = (int)(ev_slot - c->ev_step);
                for (int j = 0; j < jobs.n; j++) {
                    if (defer->n >= RED_JOBS) return RPMD_ERR_ARG;
                    defer->add(jobs.parts[j], jobs.nparts[j], base + jobs.out[j], jobs.scale[j]);
                }
                jobs.n = 0;
                kick->red = defer;
                kick->ev_out = c->ev_step;
            }
            launch_chain_rule(grad, chain_fp, chain_npart, na, nb, P.alpha, chain_scale, gacc, s, c->goo,
                              fuse ? kick : nullptr);
            if (fuse) *kick_done = defer ? 2 : 1;
        } else if (!need_oo) {
            launch_chain_rule(grad, chain_fp, chain_npart, na, nb, P.alpha, chain_scale, gacc, s);
        }
        gacc = 1;
    }
    if (need_oo && oo_branch) {
        // its job was added with the chain rule
    } else if (need_oo) {
        double* po = c->parts + c->off_oo;
        const double* chp = chain_npart ? chain_fp : nullptr;
        if (mixed)
            launch_oo<float>(pos, na, nb, bx, c->nnb, c->nbl, use_list, P.rcut, P, grad, po, s, chp, chain_npart,
                             P.alpha, chain_scale, chain_acc, oo_dmax);
        else
            launch_oo<double>(pos, na, nb, bx, c->nnb, c->nbl, use_list, P.rcut, P, grad, po, s, chp, chain_npart,
                              P.alpha, chain_scale, chain_acc, oo_dmax);
        jobs.add(po, nb * oo_blocks(na), 20, 1.0 / nb);
    }
    if (iopt == 0 || iopt == 4) {
        double* pi = c->parts + c->off_intra;
        launch_intra(pos, na, nb, intra_model(P), P, grad, pi, gacc, s);
        gacc = 1;
        jobs.add(pi, nb * intra_blocks(na), 30, 1.0 / nb);
    }
    if (iopt == 5) {
        // TEM correction (src/md_forces.f90:65-100; src/correct_tem.f90:1-71)
        if (P.wm > 0.0 || P.wh > 0.0) return RPMD_ERR_SMEARED;
        const double aeff = fabs(P.alpha - 1.0) > 1e-4 ? P.alpha : 1.0;
        int tl = 0;
        if (nb > 1) {
            double cut = P.sig + 5.0;
            if (cut > 0.5 * boxmin) cut = 0.5 * boxmin;
            launch_centroid(pos, c->cen, na, nb, s);
            launch_zero_int(c->ovf, 2, s);
            launch_nlist(c->cen, na, 1, bx, cut, c->nnb, c->nbl, c->ovf, s);
            tl = 1;
            if (check_lists) {
                int h[2];
                HIPCHK(hipMemcpyAsync(h, c->ovf, sizeof h, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (h[0] > 0) return RPMD_ERR_NEIGH_OVERFLOW;
                if (h[1] == 0) tl = 0;
            }
        }
        if ((tl || 3.0 * P.sig > boxmax) && P.sig > 0.5 * boxmax) return RPMD_ERR_EWC_CUTOFF;
        launch_sites(pos, c->site, na, nb, aeff, 0, bx, s);
        double* pt = c->parts + c->off_tem;
        launch_tem(c->site, c->q, na, nb, bx, c->nnb, c->nbl, tl, P.sig, c->gsite, pt, s);
        launch_chain_rule(grad, c->gsite, 1, na, nb, aeff, 1.0, gacc, s);
        gacc = 1;
        jobs.add(pt, nb * tem_blocks(na), 40, 1.0 / nb);
    }
    if (defer) {
        const int base = (int)(ev_slot - c->ev_step);
        for (int j = 0; j < jobs.n; j++) {
            if (defer->n >= RED_JOBS) return RPMD_ERR_ARG;
            defer->add(jobs.parts[j], jobs.nparts[j], base + jobs.out[j], jobs.scale[j]);
        }
    } else {
        launch_reduce_ev(jobs, ev_slot, s);
    }
    if (hipGetLastError() != hipSuccess) return RPMD_ERR_HIP;
    return 0;
}

// combine an ev slot read back to the host: v and vir (column-major) of one forces call
// components of an ev slot a forces call with this iopt writes: 0 real, 1 k space, 2 O-O, 3 intra, 4 TEM
bool ev_component(int iopt, int t) {
    switch (t) {
        case 0:
        case 1: return iopt <= 2;
        case 2: return iopt == 3 || iopt <= 1;
        case 3: return iopt == 0 || iopt == 4;
        default: return iopt == 5;
    }
}

void combine_ev(const rpmd_ctx* c, const double* h, int iopt, const double box[3], double* v, double* vir) {
    double vv = 0.0;
    double w[9] = {0};
    for (int t = 0; t < 5; t++) {
        if (!ev_component(iopt, t)) continue;
        vv += h[t * 10];
        for (int k = 0; k < 9; k++) w[k] += h[t * 10 + 1 + k];
    }
    if (iopt <= 2) {
        // self term -alpha/sqrt(pi) sum q^2 (src/ewald_cubic.f90:524-530; src/ewald_noncubic.f90:449-454)
        const double sc = c->prm.iamcub ? 1.0 / box[0] : 1.0;
        vv -= (c->prm.walpha / sqrt(acos(-1.0))) * c->sumq2 * sc;
    }
    *v = vv;
    if (vir) memcpy(vir, w, sizeof w);
}

int ensure_stage(rpmd_ctx* c, size_t n) {
    if (n <= c->h_cap) return 0;
    if (c->h_stage) hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    if (hipHostMalloc((void**)&c->h_stage, n * sizeof(double)) != hipSuccess) return RPMD_ERR_NOMEM;
    c->h_cap = n;
    return 0;
}

int upload_charges(rpmd_ctx* c, const double* z) {
    const int na = c->na;
    bool same = (int)c->hq.size() == na && !memcmp(c->hq.data(), z, sizeof(double) * na);
    if (!same) {
        c->hq.assign(z, z + na);
        c->sumq2 = 0.0;
        for (int i = 0; i < na; i++) c->sumq2 = c->sumq2 + z[i] * z[i];
        HIPCHK(hipMemcpyAsync(c->q, z, sizeof(double) * na, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    return 0;
}

// host AoS -> device SoA (through pinned staging)
int upload_aos(rpmd_ctx* c, const double* h, double* d_soa, int nb) {
    const size_t N = (size_t)3 * c->na * nb;
    RETCHK(ensure_stage(c, N));
    memcpy(c->h_stage, h, N * sizeof(double));
    HIPCHK(hipMemcpyAsync(c->aos, c->h_stage, N * sizeof(double), hipMemcpyHostToDevice, c->st));
    // the staging buffer is reused by the next upload: the copy must have read it first
    HIPCHK(hipStreamSynchronize(c->st));
    launch_aos_to_soa(c->aos, d_soa, c->na, nb, c->st);
    return 0;
}
int download_aos(rpmd_ctx* c, const double* d_soa, double* h, int nb) {
    const size_t N = (size_t)3 * c->na * nb;
    RETCHK(ensure_stage(c, N));
    launch_soa_to_aos(d_soa, c->aos, c->na, nb, c->st);
    HIPCHK(hipMemcpyAsync(c->h_stage, c->aos, N * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    memcpy(h, c->h_stage, N * sizeof(double));
    return 0;
}

int read_ev(rpmd_ctx* c, const double* d_slot, int iopt, const double box[3], double* v, double* vir) {
    double h[50];
    HIPCHK(hipMemcpyAsync(h, d_slot, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    combine_ev(c, h, iopt, box, v, vir);
    return 0;
}

// rp_contract_nm (src/evolve_pi.f90:324-375): rnm (nb modes) -> forces on nbdf contracted beads ->
// out_nm (nb modes, zero outside the kept modes)
int rp_contract_dev(rpmd_ctx* c, const double* rnm, int nb, int nbdf, const double box[3], int iopt, double* out_nm,
                    double* ev_slot, bool check) {
    if ((nbdf - 1) % 2 != 0) return RPMD_ERR_NBDF_EVEN;
    if (nbdf > nb) return RPMD_ERR_NBDF_TOO_LARGE;
    const int na = c->na;
    // matrices and mode maps of (nbdf, nb) are made once, outside stream capture (the first,
    // eager step of a run does it), and stay on the device
    NMMat* m = get_nm(c, nbdf, nb);
    if (!m) return RPMD_ERR_NOMEM;
    const int* dmi = m->maps;
    const int* dmo = m->maps + nbdf;
    launch_mode_copy(c->tmp2, rnm, dmi, nbdf, na, 1.0, c->st);                 // ring_contract :417-425
    launch_nm(c->tmp2, c->rb, m->Bs, nbdf, (size_t)3 * na, c->st);           // realft(-1) * sqrt(nbdf/nb)
    RETCHK(forces_dev(c, c->rb, nbdf, box, iopt, c->dvdrb, ev_slot, check));  // :366
    launch_nm(c->dvdrb, c->tmp2, m->Fs, nbdf, (size_t)3 * na, c->st);        // force_contract :384-386
    launch_mode_copy(out_nm, c->tmp2, dmo, nb, na, 1.0, c->st);               // :388-401
    return 0;
}

// thermostat coefficients for PILE-L (src/module_gle.f90:22-41) at time step dt
void pile_coeffs(const rpmd_params& P, const double* mass3, double beta, int nb, double dt, double* c1, double* c3) {
    const double hbar = 1.0, betan = beta / nb, pi = acos(-1.0), twown = 2.0 / (betan * hbar), pibyn = pi / nb;
    const double radtwo = sqrt(2.0);
    double ck[3];
    for (int j = 0; j < 3; j++) ck[j] = sqrt(mass3[j]);
    for (int k = 1; k <= nb; k++) {
        double wk = twown * sin((k - 1) * pibyn), gamma;
        if (k == 1) gamma = 2.0 * pi / P.ttau;
        else gamma = 2.0 * wk;
        double cc1 = exp(-gamma * dt);
        double c2 = sqrt((1.0 - cc1 * cc1) * 0.5 / betan);
        if (k == 1 || ((nb % 2 == 0) && k == (nb / 2 + 1))) c2 = radtwo * c2;
        c1[k - 1] = cc1;
        for (int j = 0; j < 3; j++) c3[(k - 1) * 3 + j] = c2 * ck[j];
    }
}

// ring_rpmd (src/md_pathi.f90:190-231): monod(3,4,nb)
void ring_rpmd(const double* mass3, double* monod, int nb, double dt, double beta) {
    const double hbar = 1.0, pi = acos(-1.0);
    memset(monod, 0, sizeof(double) * 12 * nb);
    auto MON = [&](int j, int a, int k) -> double& { return monod[((size_t)k * 4 + a) * 3 + j]; };
    for (int j = 0; j < 3; j++) {
        MON(j, 0, 0) = 1.0;
        MON(j, 1, 0) = 0.0;
        MON(j, 2, 0) = dt / mass3[j];
        MON(j, 3, 0) = 1.0;
        if (nb > 1) {
            double betan = beta / (double)nb, twown = 2.0 / (betan * hbar), pibyn = pi / (double)nb;
            for (int k = 1; k <= nb / 2; k++) {
                double wk = twown * sin(k * pibyn), wt = wk * dt, wm = wk * mass3[j];
                double cwt = cos(wt), swt = sin(wt);
                MON(j, 0, k) = cwt;
                MON(j, 1, k) = -wm * swt;
                MON(j, 2, k) = swt / wm;
                MON(j, 3, k) = cwt;
            }
            for (int k = 1; k <= (nb - 1) / 2; k++)
                for (int a = 0; a < 4; a++) MON(j, a, nb - k) = MON(j, a, k);
        }
    }
}

// GLE propagators (src/module_gle.f90:156-195).  A is (n x n) in the GLE-A file's row order, in
// units of 2 pi / ttau, where the reference forms 2 pi in single precision (2.*3.1415927).
// T = exp(-dt A) by matrix_exp(.., 15, 15) (:323-354: Taylor series of 15 terms by Horner's rule
// on A / 2^15, then 15 squarings); S = cholesky(C - T C T^T) (:357-378) with C = (nb/beta) I.
// T and S are returned column-major.
void gle_propagators(const double* A, int n, double ttau, int nb, double dt, double beta, double* T, double* S) {
    auto at = [n](double* m, int i, int j) -> double& { return m[(size_t)j * n + i]; };
    std::vector<double> M((size_t)n * n), E((size_t)n * n), W((size_t)n * n), C((size_t)n * n, 0.0);
    const float twopi_f = 2.0f * 3.1415927f;
    const double fac = (double)twopi_f / ttau;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) at(M.data(), i, j) = -dt * (A[(size_t)i * n + j] * fac);
    const int jt = 15, ks = 15;
    std::vector<double> tc(jt + 1);
    tc[0] = 1.0;
    for (int i = 1; i <= jt; i++) tc[i] = tc[i - 1] / (double)i;
    const double scl = (double)(1.0f / powf(2.0f, (float)ks));
    for (double& x : M) x = x * scl;
    auto matmul = [&](const std::vector<double>& X, const std::vector<double>& Y, std::vector<double>& Z) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double acc = 0.0;
                for (int c = 0; c < n; c++) acc += X[(size_t)c * n + i] * Y[(size_t)j * n + c];
                Z[(size_t)j * n + i] = acc;
            }
    };
    std::fill(E.begin(), E.end(), 0.0);
    for (int i = 0; i < n; i++) at(E.data(), i, i) = tc[jt];
    for (int q = jt; q >= 1; q--) {
        matmul(M, E, W);
        E.swap(W);
        for (int i = 0; i < n; i++) at(E.data(), i, i) += tc[q - 1];
    }
    for (int q = 0; q < ks; q++) {
        matmul(E, E, W);
        E.swap(W);
    }
    for (int i = 0; i < n; i++) at(C.data(), i, i) = (double)nb / beta;
    // D = C - T C T^T
    std::vector<double> TC((size_t)n * n), TCT((size_t)n * n), Et((size_t)n * n), D((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Et[(size_t)j * n + i] = E[(size_t)i * n + j];
    matmul(C, Et, TC);
    matmul(E, TC, TCT);
    for (size_t i = 0; i < D.size(); i++) D[i] = C[i] - TCT[i];
    std::vector<double> L((size_t)n * n, 0.0);
    at(L.data(), 0, 0) = sqrt(at(D.data(), 0, 0));
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < i; j++) {
            double x = at(D.data(), i, j);
            for (int k = 0; k < j; k++) x = x - at(L.data(), i, k) * at(L.data(), j, k);
            at(L.data(), i, j) = x / at(L.data(), j, j);
        }
        double x = at(D.data(), i, i);
        for (int k = 0; k < i; k++) x = x - at(L.data(), i, k) * at(L.data(), i, k);
        at(L.data(), i, i) = sqrt(x);
    }
    memcpy(T, E.data(), sizeof(double) * n * n);
    memcpy(S, L.data(), sizeof(double) * n * n);
}

// noise source for one draw block of 3*na*nb deviates
struct NoiseCursor {
    const double* base;  // device pointer of host-supplied deviates, or null (device stream)
    size_t used;
    size_t cap;
    uint64_t stream_add;
    const double* take(size_t n) {
        if (!base) return nullptr;
        const double* p = base + used;
        used += n;
        return p;
    }
};

// one evolve_pi step (type 'RPMD', nbaro = 0) on the context's device state
// (src/evolve_pi.f90:1-322).  Energies -> c->ev_step slots; vir_hf averaged on the host.
// full_forces (src/md_forces.f90:305-420) on device positions pos (SoA) with the charges uploaded:
// the high-frequency gradient into gh, the low-frequency one into *gl (a pointer to one of the two
// buffers gl, gt: the last normal-mode transform may leave it in either), with go, ge as scratch
// (and rnm for the normal modes); energies and the virial on the host
int full_forces_core(rpmd_ctx* c, const double* pos, int nb, const double box[3], double** gl_io, double* gh, double* go,
                     double* ge, double* gt, double* v, double* vew, double* voo, double* vint, double* vir) {
    const rpmd_params& P = c->prm;
    const int na = c->na;
    const size_t N = (size_t)3 * na * nb;
    hipStream_t s = c->st;
    double* gl = *gl_io;
    double e_int, w_int[9], e_ew = 0, w_ew[9] = {0}, e_oo = 0, w_oo[9] = {0}, tv, tw[9];
    // high-frequency forces (src/md_forces.f90:350)
    RETCHK(forces_dev(c, pos, nb, box, 4, gh, c->ev, true));
    RETCHK(read_ev(c, c->ev, 4, box, &e_int, w_int));
    launch_zero(gl, N, s);
    launch_zero(go, N, s);
    const int nbdf1 = P.nbdf1, nbdf2 = P.nbdf2;
    if (nb > 1 && (nbdf1 > 0 || nbdf2 > 0)) {
        NMMat* m = get_nm(c, nb, 0);
        if (!m) return RPMD_ERR_NOMEM;
        launch_nm(pos, c->rnm, m->F, nb, (size_t)3 * na, s);  // :359
        if (nbdf1 > 0) {
            RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf1, box, 2, gl, c->ev, true));
            RETCHK(read_ev(c, c->ev, 2, box, &e_ew, w_ew));
            if (P.sig > 0.0) {
                RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf1, box, 5, ge, c->ev, true));
                RETCHK(read_ev(c, c->ev, 5, box, &tv, tw));
                e_ew -= tv;
                for (int k = 0; k < 9; k++) w_ew[k] -= tw[k];
                launch_axpy(gl, ge, -1.0, N, s);
            }
        }
        if (nbdf2 > 0) {
            RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf2, box, 3, go, c->ev, true));
            RETCHK(read_ev(c, c->ev, 3, box, &e_oo, w_oo));
        }
        if (nbdf1 > 0 && nbdf2 > 0) {
            launch_axpy(gl, go, 1.0, N, s);
            launch_nm(gl, gt, m->B, nb, (size_t)3 * na, s);
            std::swap(gl, gt);
        } else if (nbdf1 > 0) {
            launch_nm(gl, gt, m->B, nb, (size_t)3 * na, s);
            std::swap(gl, gt);
        } else if (nbdf2 > 0) {
            launch_nm(go, gt, m->B, nb, (size_t)3 * na, s);
            std::swap(go, gt);
        }
    }
    if (nb == 1 || nbdf1 == 0) {
        RETCHK(forces_dev(c, pos, nb, box, 2, gl, c->ev, true));
        RETCHK(read_ev(c, c->ev, 2, box, &e_ew, w_ew));
    } else if (P.sig > 0.0) {
        RETCHK(forces_dev(c, pos, nb, box, 5, ge, c->ev, true));
        RETCHK(read_ev(c, c->ev, 5, box, &tv, tw));
        e_ew += tv;
        for (int k = 0; k < 9; k++) w_ew[k] += tw[k];
        launch_axpy(gl, ge, 1.0, N, s);
    }
    if (nb == 1 || nbdf2 == 0) {
        RETCHK(forces_dev(c, pos, nb, box, 3, go, c->ev, true));
        RETCHK(read_ev(c, c->ev, 3, box, &e_oo, w_oo));
    }
    if (nb == 1 || nbdf1 <= 0 || nbdf2 <= 0) launch_axpy(gl, go, 1.0, N, s);
    *gl_io = gl;
    if (vew) *vew = e_ew;
    if (voo) *voo = e_oo;
    if (vint) *vint = e_int;
    if (v) *v = e_ew + e_oo + e_int;
    if (vir)
        for (int k = 0; k < 9; k++) vir[k] = w_oo[k] + w_ew[k] + w_int[k];
    return 0;
}

// one barostat step's inputs (evolve_pi with nbaro = 1): the previous step's low-frequency virial and
// the MC draws, replayed in order (u) or, when u is null, from the library's own stream
struct BaroCall {
    double vir_lf[9] = {0};
    const double* u = nullptr;
    int nu = 0, ui = 0;
};

// a ran2(irun, lo, hi) draw of the barostat: the next replayed value, or lo + (hi - lo) x with x a
// 53-bit uniform of splitmix64 over the context seed and a per-context draw counter
int baro_uniform(rpmd_ctx* c, BaroCall* bc, double lo, double hi, double* out) {
    if (bc->u) {
        if (bc->ui >= bc->nu) return RPMD_ERR_NOISE;
        *out = bc->u[bc->ui++];
        return 0;
    }
    uint64_t z = c->seed * 0x9E3779B97F4A7C15ull + (++c->baro_draws) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    *out = lo + (hi - lo) * ((double)(z >> 11) * 0x1p-53);
    return 0;
}

// the buffers of the barostat's virial_ke sums (BER)
int baro_buffers(rpmd_ctx* c, int mts) {
    if (!c->baro_parts) RETCHK(dalloc(&c->baro_parts, (size_t)estimator_blocks(c->na) * 10));
    if (c->baro_mts < mts) {
        if (c->baro_ev) hipFree(c->baro_ev);
        c->baro_ev = nullptr;
        RETCHK(dalloc(&c->baro_ev, (size_t)mts * 10));
        c->baro_mts = mts;
    }
    return 0;
}

// pres_lj_tail (src/md_pressure.f90:37-56): the LJ tail correction to the pressure
double pres_lj_tail(const rpmd_params& P, const double box[3], int na) {
    const int nm = na / 3;
    const double pi = acos(-1.0), vol = box[0] * box[1] * box[2], rho = (double)nm / vol;
    const double prefac = (16.0 * pi * (rho * rho) * P.oo_eps * (P.oo_sig * P.oo_sig * P.oo_sig)) / 3.0;
    const double x = P.oo_sig / P.rcut, x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x8 = x4 * x4, x9 = x8 * x;
    return prefac * ((2.0 / 3.0) * x9 - x3);
}

// the barostat of evolve_pi with nbaro = 1 (src/evolve_pi.f90:163-189) on the bead positions rst
// after the multiple time steps; the box of the rest of the step (st_box) follows it, and the
// normal-mode positions rnm become those of the moved positions (r = rst; realft(r, +1)).
// BER (src/module_barostat.f90:8-122): the pressure of vir = vir_lf + the substeps' mean
// high-frequency virial and of the substeps' mean virial_ke, isotropic in a cubic box, per axis
// otherwise (with the module's patm: the reference's beren_baro_nc hides it behind a local of the
// same name, :34), compressibility 13531 a.u.; MCI (:124-220): with probability 1/taub a volume
// move vn = vo + dv u(-1, 1) of the molecules, judged on full_forces energies at both volumes.
// Scratch while it runs: p (rebuilt from pnm after), pos, grad, dvdrl, dvdre, tmp, rnm.
int baro_step(rpmd_ctx* c, BaroCall* bc, int nb, int mts) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nm = na / 3;
    const size_t N = (size_t)3 * na * nb;
    hipStream_t s = c->st;
    double box[3] = {c->st_box.L[0], c->st_box.L[1], c->st_box.L[2]};
    const double beta = c->st_beta, dt = c->st_dt;
    if (c->baro == RPMD_BARO_BER) {
        std::vector<double> h((size_t)mts * 50), t((size_t)mts * 10);
        HIPCHK(hipMemcpyAsync(h.data(), c->ev_step, h.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(t.data(), c->baro_ev, t.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        // virial_ke of each substep (src/kinetic.f90:42-49; nc = 3, nctot = 0), its mean, and the mean
        // high-frequency virial (src/evolve_pi.f90:146-151, 163-166)
        const double tc = 0.5 * (double)(3 * na - 3) / beta;
        double tv = 0.0, tvxyz[3] = {0, 0, 0}, whf[9] = {0};
        for (int im = 0; im < mts; im++) {
            const double* q = &t[(size_t)im * 10];
            const double tq1 = 0.5 * (q[2] + q[3] + q[4]) / nb, tq2 = 0.5 * (q[5] + q[6] + q[7]) / nb;
            tv = tv + (tc + tq1 + tq2);
            for (int d = 0; d < 3; d++) tvxyz[d] = tvxyz[d] + (tc / 3.0 + 0.5 * q[2 + d] / nb + 0.5 * q[5 + d] / nb);
            double vv, w[9];
            combine_ev(c, &h[(size_t)im * 50], 4, box, &vv, w);
            for (int k = 0; k < 9; k++) whf[k] += w[k];
        }
        tv = tv / mts;
        double vir[9];
        for (int d = 0; d < 3; d++) tvxyz[d] = tvxyz[d] / mts;
        for (int k = 0; k < 9; k++) vir[k] = bc->vir_lf[k] + whf[k] / (double)mts;
        const double bstat = 13531.0 / (3.0 * c->taub), vol = box[0] * box[1] * box[2];
        const double ptail = P.oo_gam == 0.0 ? pres_lj_tail(P, box, na) : 0.0;
        double scale[3];
        if (P.iamcub) {
            // pressure (src/md_pressure.f90:1-35)
            const double w = -(vir[0] + vir[4] + vir[8]) / 3.0;
            const double pres = (2.0 * tv / 3.0 + w) / vol + ptail;
            scale[0] = scale[1] = scale[2] = fabs(1.0 - (dt * bstat) * (c->patm - pres));
        } else {
            for (int d = 0; d < 3; d++) {
                const double pres = (-vir[4 * d] + 2.0 * tvxyz[d]) / vol + ptail;
                scale[d] = fabs(1.0 - (dt * bstat) * (c->patm - pres));
            }
        }
        for (int d = 0; d < 3; d++) box[d] = scale[d] * box[d];
        launch_baro_shift(c->rst, na, nb, scale, s);
    } else if (c->baro == RPMD_BARO_MCI) {
        double u;
        RETCHK(baro_uniform(c, bc, 0.0, 1.0, &u));
        if (u < 1.0 / c->taub) {
            double eo, en, *gl = c->grad;
            RETCHK(full_forces_core(c, c->rst, nb, box, &gl, c->p, c->dvdrl, c->dvdre, c->tmp, &eo, nullptr, nullptr,
                                    nullptr, nullptr));
            c->nmove++;
            const double vo = box[0] * box[1] * box[2];
            double u2, u3;
            RETCHK(baro_uniform(c, bc, -1.0, 1.0, &u2));
            const double vn = vo + c->dv * u2;
            const double sc = pow(vn / vo, 1.0 / 3.0);
            const double boxn[3] = {sc * box[0], sc * box[1], sc * box[2]}, scale[3] = {sc, sc, sc};
            HIPCHK(hipMemcpyAsync(c->pos, c->rst, N * 8, hipMemcpyDeviceToDevice, s));
            launch_baro_shift(c->pos, na, nb, scale, s);
            gl = c->grad;
            RETCHK(full_forces_core(c, c->pos, nb, boxn, &gl, c->p, c->dvdrl, c->dvdre, c->tmp, &en, nullptr, nullptr,
                                    nullptr, nullptr));
            double test = (en - eo) + c->patm * (vn - vo) - (double)nm * log(vn / vo) / beta;
            test = exp(-beta * test);
            RETCHK(baro_uniform(c, bc, 0.0, 1.0, &u3));
            if (test > u3) {
                // accepted: the box, the positions and their high-frequency forces
                c->nacc++;
                for (int d = 0; d < 3; d++) box[d] = boxn[d];
                HIPCHK(hipMemcpyAsync(c->rst, c->pos, N * 8, hipMemcpyDeviceToDevice, s));
                HIPCHK(hipMemcpyAsync(c->dvdr2, c->p, N * 8, hipMemcpyDeviceToDevice, s));
            }
        }
    } else {
        return RPMD_ERR_BARO;
    }
    c->st_box = make_box(box);
    NMMat* m = get_nm(c, nb, 0);
    if (!m) return RPMD_ERR_NOMEM;
    launch_nm(c->rst, c->rnm, m->F, nb, (size_t)3 * na, s);
    return 0;
}

// the barostat of evolve_rig_pi with nbaro = 1 (src/evolve_rigid.f90:203-220) after the rigid step:
// dvdr2 is zero; BER from virial_ke of the final positions with the step's intermolecular gradient
// (nctot = 3 nm) and the step's in/out virial vir; MCI replaces v, vir, dvdr and dvdr2 by its
// full_forces values when it tries a move, and by those of the moved molecules when it accepts.
// Scratch: pos, grad, pnm, dvdrl, dvdre, tmp, rnm.
int rig_baro_tail(rpmd_ctx* c, BaroCall* bc, int nb, double* v, double vir[9]) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nm = na / 3;
    const size_t N = (size_t)3 * na * nb;
    hipStream_t s = c->st;
    double box[3] = {c->st_box.L[0], c->st_box.L[1], c->st_box.L[2]};
    const double beta = c->st_beta, dt = c->st_dt;
    if (c->baro == RPMD_BARO_BER) {
        launch_estimators(c->p, c->r, c->dvdr, c->dvdr2, c->mass, na, nb, 0.0, c->baro_parts, s);
        ReduceJobs bj;
        bj.add(c->baro_parts, estimator_blocks(na), 0, 1.0);
        launch_reduce_ev(bj, c->baro_ev, s);
        double q[10];
        HIPCHK(hipMemcpyAsync(q, c->baro_ev, sizeof q, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const double tc = 0.5 * (double)(3 * na - 3 - 3 * nm) / beta;
        const double tv = tc + 0.5 * (q[2] + q[3] + q[4]) / nb + 0.5 * (q[5] + q[6] + q[7]) / nb;
        double tvxyz[3];
        for (int d = 0; d < 3; d++) tvxyz[d] = tc / 3.0 + 0.5 * q[2 + d] / nb + 0.5 * q[5 + d] / nb;
        const double bstat = 13531.0 / (3.0 * c->taub), vol = box[0] * box[1] * box[2];
        const double ptail = P.oo_gam == 0.0 ? pres_lj_tail(P, box, na) : 0.0;
        double scale[3];
        if (P.iamcub) {
            const double w = -(vir[0] + vir[4] + vir[8]) / 3.0;
            const double pres = (2.0 * tv / 3.0 + w) / vol + ptail;
            scale[0] = scale[1] = scale[2] = fabs(1.0 - (dt * bstat) * (c->patm - pres));
        } else {
            for (int d = 0; d < 3; d++) {
                const double pres = (-vir[4 * d] + 2.0 * tvxyz[d]) / vol + ptail;
                scale[d] = fabs(1.0 - (dt * bstat) * (c->patm - pres));
            }
        }
        for (int d = 0; d < 3; d++) box[d] = scale[d] * box[d];
        launch_baro_shift(c->r, na, nb, scale, s);
    } else if (c->baro == RPMD_BARO_MCI) {
        double u;
        RETCHK(baro_uniform(c, bc, 0.0, 1.0, &u));
        if (u < 1.0 / c->taub) {
            double eo, en, wo[9], wn[9], *gl = c->grad;
            RETCHK(full_forces_core(c, c->r, nb, box, &gl, c->pnm, c->dvdrl, c->dvdre, c->tmp, &eo, nullptr, nullptr,
                                    nullptr, wo));
            // mc_baro keeps the forces of its first full_forces call (src/module_barostat.f90:147-148)
            HIPCHK(hipMemcpyAsync(c->dvdr, gl, N * 8, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(c->dvdr2, c->pnm, N * 8, hipMemcpyDeviceToDevice, s));
            *v = eo;
            for (int k = 0; k < 9; k++) vir[k] = wo[k];
            c->nmove++;
            const double vo = box[0] * box[1] * box[2];
            double u2, u3;
            RETCHK(baro_uniform(c, bc, -1.0, 1.0, &u2));
            const double vn = vo + c->dv * u2;
            const double sc = pow(vn / vo, 1.0 / 3.0);
            const double boxn[3] = {sc * box[0], sc * box[1], sc * box[2]}, scale[3] = {sc, sc, sc};
            HIPCHK(hipMemcpyAsync(c->pos, c->r, N * 8, hipMemcpyDeviceToDevice, s));
            launch_baro_shift(c->pos, na, nb, scale, s);
            gl = c->grad;
            RETCHK(full_forces_core(c, c->pos, nb, boxn, &gl, c->pnm, c->dvdrl, c->dvdre, c->tmp, &en, nullptr, nullptr,
                                    nullptr, wn));
            double test = (en - eo) + c->patm * (vn - vo) - (double)nm * log(vn / vo) / beta;
            test = exp(-beta * test);
            RETCHK(baro_uniform(c, bc, 0.0, 1.0, &u3));
            if (test > u3) {
                c->nacc++;
                for (int d = 0; d < 3; d++) box[d] = boxn[d];
                HIPCHK(hipMemcpyAsync(c->r, c->pos, N * 8, hipMemcpyDeviceToDevice, s));
                HIPCHK(hipMemcpyAsync(c->dvdr, gl, N * 8, hipMemcpyDeviceToDevice, s));
                HIPCHK(hipMemcpyAsync(c->dvdr2, c->pnm, N * 8, hipMemcpyDeviceToDevice, s));
                *v = en;
                for (int k = 0; k < 9; k++) vir[k] = wn[k];
            }
        }
    } else {
        return RPMD_ERR_BARO;
    }
    c->st_box = make_box(box);
    return 0;
}

int step_dev(rpmd_ctx* c, NoiseCursor& nz, bool check, BaroCall* bc) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nb = c->st_nb, mts = P.mts;
    const size_t N = (size_t)3 * na * nb, ne = (size_t)3 * na;
    hipStream_t s = c->st;
    const double dt = c->st_dt, dtsmall = dt / (double)mts, halfdtsmall = 0.5 * dtsmall, halfdt = 0.5 * dt;
    const double* box = c->st_box.L;
    NMMat* nmm = get_nm(c, nb, 0);
    if (!nmm) return RPMD_ERR_NOMEM;
    const size_t draws = (size_t)3 * na * nb;
    if (P.therm == RPMD_THERM_GLE && (!c->gle_gp || c->gle_nb != nb)) return RPMD_ERR_GLE;
    const double beta = c->st_beta, betan = beta / (double)nb;
    auto outer = [&]() {
        if (P.therm == RPMD_THERM_PRL) {
            // parinello_therm_loc with dt (src/evolve_pi.f90:75-77, src/module_thermostat.f90:56-82)
            const double gamma = 1.0 / (2.0 * P.ttau), c1 = exp(-0.5 * gamma * dt);
            launch_langevin(c->p, c->lang_c3m, c1, na, nb, nz.take(draws), c->seed, c->step_ctr, nz.stream_add++, s);
        } else if (P.therm == RPMD_THERM_PRG) {
            // parinello_therm(p, mass, ttau, na, nb, 0.5 dt, irun, beta) on the beads, nctot = 0
            const long long ndof = 3LL * na * nb - 3;
            launch_bussi(c->p, c->mass, na, nb, ndof, betan, exp(-(0.5 * dt) / P.ttau), nz.take((size_t)ndof + 1),
                         c->seed, c->step_ctr, nz.stream_add++, c->bussi_parts, c->bussi_alpha, nullptr, s);
        } else if (P.therm == RPMD_THERM_GLE) {
            launch_gle(c->p, c->gle_gp, c->gle_TS, c->mass, na, nb, c->gle_n1, nz.take(draws * c->gle_n1), c->seed,
                       c->step_ctr, nz.stream_add++, c->gle_parts, c->gle_pcom, 1.0 / sqrt(nb * c->mtot), s);
        }
    };
    const int nblk = pile_blocks(na);
    const double inv_mtot = 1.0 / c->mtot;
    if (P.therm == RPMD_THERM_PRO && mts == 1 && P.nbdf1 <= 0 && P.nbdf2 <= 0 && !bc) {
        // the same sequence with fewer launches: the opening kick is applied as p is loaded by the NM
        // transform (p itself is recomputed from pnm below), each COM removal as pnm is loaded by the
        // next kernel, the free ring and the second thermostat kick are one launch, p = B pnm is taken
        // before the intramolecular forces (which do not read it), the two closing kicks are one launch,
        // and the noise counter advances in it
        const double* n1 = nz.take(draws);
        const double* n2 = nz.take(draws);
        const uint64_t sa1 = nz.stream_add++, sa2 = nz.stream_add++;
        // up to 64 beads the whole sequence is one wave per 16 columns plus the COM corrections
        if (!launch_pile_step(c->p, c->dvdr, halfdt, c->dvdr2, halfdtsmall, c->r, c->p, c->rst, nmm->F, nmm->B, nb,
                              na, c->monod, c->mass, inv_mtot, c->pile_c1, c->pile_c3, n1, n2, c->seed, c->step_ctr,
                              sa1, c->pstep_parts, s)) {
            launch_nm_kick2(c->p, c->dvdr, halfdt, c->dvdr2, halfdtsmall, c->pnm, c->r, c->rnm, nmm->F, nb, ne, s);
            launch_pile(c->pnm, c->pile_c1, c->pile_c3, na, nb, 0, n1, c->seed, c->step_ctr, sa1, c->com_parts, s);
            launch_free_ring_pile(c->pnm, c->rnm, c->monod, na, nb, c->com_parts, inv_mtot, c->mass, c->pile_c1,
                                  c->pile_c3, n2, c->seed, c->step_ctr, sa2, c->com_parts2, s);
            launch_nm_com2(c->rnm, c->rst, c->pnm, c->p, c->com_parts2, nblk, inv_mtot, c->mass, na, nmm->B, nb, ne,
                           s);
        }
        ReduceJobs red;
        // the intramolecular forces on the third stream, beside the other force kernels; on two
        // streams they follow the O-O terms on the second (the same positions: rst becomes r)
        const bool intra3 = c->streams > 2 && !check && !c->prof;
        const bool intra2 = c->streams == 2 && !check && !c->prof;
        if (intra3) {
            HIPCHK(hipEventRecord(c->ev_fork3, s));
            HIPCHK(hipStreamWaitEvent(c->st3, c->ev_fork3, 0));
        }
        if (!intra2)
            RETCHK(forces_dev(c, c->rst, nb, box, 4, c->dvdr2, c->ev_step, check, &red, intra3 ? c->st3 : nullptr));
        std::swap(c->r, c->rst);
        c->lf_merged = true;
        int intra_done = 0, kick_done = 0;
        // the closing kicks fused into the chain-rule kernel when the intramolecular forces (dvdr2)
        // are done before it, i.e. on the second stream
        KickArgs kick;
        kick.p = c->p;
        kick.g1 = c->dvdr2;
        kick.a1 = halfdtsmall;
        kick.a2 = halfdt;
        kick.ctr = c->step_ctr;
        RETCHK(forces_dev(c, c->r, nb, box, 1, c->dvdr, c->ev_step + (size_t)mts * 50, check, &red, nullptr,
                          intra2 ? c->dvdr2 : nullptr, &intra_done, intra2 ? &kick : nullptr, &kick_done));
        if (intra2 && !intra_done) RETCHK(forces_dev(c, c->r, nb, box, 4, c->dvdr2, c->ev_step, check, &red));
        if (intra3) {
            HIPCHK(hipEventRecord(c->ev_join3, c->st3));
            HIPCHK(hipStreamWaitEvent(s, c->ev_join3, 0));
        }
        if (!(intra_done && kick_done)) launch_kick(c->p, c->dvdr2, halfdtsmall, c->dvdr, halfdt, N, s, c->step_ctr);
        // kick_done == 2: the reductions ran with the chain rule (every job was queued by then)
        if (!(intra_done && kick_done == 2)) launch_reduce_ev(red, c->ev_step, s);
        if (nz.stream_add > RPMD_STREAMS_PER_STEP) return RPMD_ERR_ARG;
        if (hipGetLastError() != hipSuccess) return RPMD_ERR_HIP;
        return 0;
    }
    outer();
    launch_kick(c->p, c->dvdr, halfdt, c->dvdr2, halfdtsmall, N, s);
    launch_nm(c->p, c->pnm, nmm->F, nb, ne, s);
    launch_nm(c->r, c->rnm, nmm->F, nb, ne, s);
    auto pile = [&]() {
        if (P.therm == RPMD_THERM_PRO) {
            launch_pile(c->pnm, c->pile_c1, c->pile_c3, na, nb, 0, nz.take(draws), c->seed, c->step_ctr,
                        nz.stream_add++, c->com_parts, s);
            launch_com_remove(c->pnm, c->mass, c->com_parts, nblk, inv_mtot, na, s);
        } else if (P.therm == RPMD_THERM_PRM) {
            // therm_pro_global (src/module_gle.f90:75-143): PILE on modes 2..nb, Bussi on the
            // centroid (parinello_therm with nb = 1 and beta_n), then the centroid COM removal
            launch_pile(c->pnm, c->pile_c1, c->pile_c3, na, nb, 1, nz.take((size_t)3 * na * (nb - 1)), c->seed,
                        c->step_ctr, nz.stream_add++, nullptr, s);
            const long long ndof = 3LL * na - 3;
            launch_bussi(c->pnm, c->mass, na, 1, ndof, betan, exp(-halfdtsmall / P.ttau), nz.take((size_t)ndof + 1),
                         c->seed, c->step_ctr, nz.stream_add++, c->bussi_parts, c->bussi_alpha, c->com_parts, s);
            launch_com_remove(c->pnm, c->mass, c->com_parts, nblk, inv_mtot, na, s);
        }
    };
    double* ev_hf = c->ev_step;
    for (int im = 1; im <= mts; im++) {
        pile();
        launch_free_ring(c->pnm, c->rnm, c->monod, na, nb, s);
        pile();
        launch_nm(c->rnm, c->rst, nmm->B, nb, ne, s);
        RETCHK(forces_dev(c, c->rst, nb, box, 4, c->dvdr2, ev_hf + (size_t)(im - 1) * 50, check));
        if (bc && c->baro == RPMD_BARO_BER) {
            // virial_ke of the substep (src/evolve_pi.f90:146-151): the centroid-virial sums of the
            // bead positions with the step's low- and the substep's high-frequency gradients
            launch_estimators(c->pnm, c->rst, c->dvdr, c->dvdr2, c->mass, na, nb, 0.0, c->baro_parts, s);
            ReduceJobs bj;
            bj.add(c->baro_parts, estimator_blocks(na), (im - 1) * 10, 1.0);
            launch_reduce_ev(bj, c->baro_ev, s);
        }
        if (im != mts) {
            launch_nm(c->dvdr2, c->tmp, nmm->F, nb, ne, s);
            launch_kick(c->pnm, c->tmp, dtsmall, nullptr, 0.0, N, s);
        }
    }
    if (bc) RETCHK(baro_step(c, bc, nb, mts));
    double* ev_lf = c->ev_step + (size_t)mts * 50;  // [EW, EWC-contracted, LJ, EWC-bead] x 50
    const int nbdf1 = P.nbdf1, nbdf2 = P.nbdf2;
    if (nbdf1 > 0) {
        RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf1, box, 2, c->dvdr, ev_lf, check));
        if (P.sig > 0.0) {
            RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf1, box, 5, c->dvdre, ev_lf + 50, check));
            launch_axpy(c->dvdr, c->dvdre, -1.0, N, s);
        }
    }
    if (nbdf2 > 0) RETCHK(rp_contract_dev(c, c->rnm, nb, nbdf2, box, 3, c->dvdrl, ev_lf + 100, check));
    launch_nm(c->pnm, c->p, nmm->B, nb, ne, s);
    std::swap(c->r, c->rst);
    launch_kick(c->p, c->dvdr2, halfdtsmall, nullptr, 0.0, N, s);
    if (nbdf1 > 0 && nbdf2 > 0) {
        launch_axpy(c->dvdr, c->dvdrl, 1.0, N, s);
        launch_nm(c->dvdr, c->tmp, nmm->B, nb, ne, s);
        std::swap(c->dvdr, c->tmp);
    } else if (nbdf1 > 0) {
        launch_nm(c->dvdr, c->tmp, nmm->B, nb, ne, s);
        std::swap(c->dvdr, c->tmp);
    } else if (nbdf2 > 0) {
        launch_nm(c->dvdrl, c->tmp, nmm->B, nb, ne, s);
        std::swap(c->dvdrl, c->tmp);
    }
    c->lf_merged = nbdf1 <= 0 && nbdf2 <= 0;
    if (c->lf_merged) {
        // Ewald then O-O into one gradient, as forces(iopt = 1) sums them (src/md_forces.f90:192-236);
        // both energies stay separate in the slot
        RETCHK(forces_dev(c, c->r, nb, box, 1, c->dvdr, ev_lf, check));
    } else {
        if (nbdf1 <= 0) {
            RETCHK(forces_dev(c, c->r, nb, box, 2, c->dvdr, ev_lf, check));
        } else if (P.sig > 0.0) {
            RETCHK(forces_dev(c, c->r, nb, box, 5, c->dvdre, ev_lf + 150, check));
            launch_axpy(c->dvdr, c->dvdre, 1.0, N, s);
        }
        if (nbdf2 <= 0) RETCHK(forces_dev(c, c->r, nb, box, 3, c->dvdrl, ev_lf + 100, check));
        if (nbdf1 <= 0 || nbdf2 <= 0) launch_axpy(c->dvdr, c->dvdrl, 1.0, N, s);
    }
    launch_kick(c->p, c->dvdr, halfdt, nullptr, 0.0, N, s);
    outer();
    launch_counter_inc(c->step_ctr, s);
    if (nz.stream_add > RPMD_STREAMS_PER_STEP) return RPMD_ERR_ARG;
    if (hipGetLastError() != hipSuccess) return RPMD_ERR_HIP;
    return 0;
}

// the RATTLE tolerance of the parameters (<= 0: the reference's 1e-6, src/md_rattle.f90:60,182)
double rattle_tol(const rpmd_params& P) { return P.rat_tol > 0.0 ? P.rat_tol : 1e-6; }

// one rigid RPMD step, evolve_rig_pi (src/evolve_rigid.f90:100-234, nbaro = 0), on the context's
// device state.  The reference's dvdr_corr, v_corr and vir_corr are never set there; they are zero
// here.  ev_step[0..49]: the forces(iopt = 1) slot; ev_step[50..58]: vir_hf (column-major).
// Thermostats PRG and PRL as the reference's; PRO and GLE are additions of this library (the
// reference's rigid step has no such branch, src/evolve_rigid.f90:123-127): at the same two points
// of the step, a PILE-L kick of the normal-mode momenta (src/module_gle.f90:8-73) or a GLE kick
// (:229-283) over dt/2, then the velocity stage of RATTLE on every bead, which puts the momenta
// back on the constraint surface.
int step_rig_dev(rpmd_ctx* c, NoiseCursor& nz, bool check) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nb = c->st_nb, nm = na / 3;
    const size_t N = (size_t)3 * na * nb;
    hipStream_t s = c->st;
    const double dt = c->st_dt, halfdt = 0.5 * dt, smalldt = dt / (double)nb, halfsmalldt = 0.5 * smalldt;
    const double beta = c->st_beta, betan = beta / (double)nb;
    const double* box = c->st_box.L;
    double omegan = (double)nb / (beta * 1.0);
    omegan = omegan * omegan;
    // rattle_setup (src/md_rattle.f90:14-28)
    const double rehh = 2.0 * sin(0.5 * P.theta) * P.reoh;
    const double len2[3] = {P.reoh * P.reoh, P.reoh * P.reoh, rehh * rehh};
    const double tol = rattle_tol(P);
    if (P.therm == RPMD_THERM_GLE && (!c->gle_gp || c->gle_nb != nb)) return RPMD_ERR_GLE;
    NMMat* nmm = P.therm == RPMD_THERM_PRO ? get_nm(c, nb, 0) : nullptr;
    if (P.therm == RPMD_THERM_PRO && !nmm) return RPMD_ERR_NOMEM;
    const size_t draws = (size_t)3 * na * nb;
    auto thermo = [&]() {
        if (P.therm == RPMD_THERM_PRG) {
            // parinello_therm with nctot = 3 nm constraints (src/md_rattle.f90:29)
            const long long ndof = 3LL * na * nb - 3LL * nm - 3;
            launch_bussi(c->p, c->mass, na, nb, ndof, betan, exp(-halfdt / P.ttau), nz.take((size_t)ndof + 1),
                         c->seed, c->step_ctr, nz.stream_add++, c->bussi_parts, c->bussi_alpha, nullptr, s);
        } else if (P.therm == RPMD_THERM_PRL) {
            const double gamma = 1.0 / (2.0 * P.ttau), c1 = exp(-0.5 * gamma * halfdt);
            launch_langevin(c->p, c->lang_c3m, c1, na, nb, nz.take(N), c->seed, c->step_ctr, nz.stream_add++, s);
        } else if (P.therm == RPMD_THERM_PRO) {
            // PILE-L over dt/2 on the normal modes (coefficients of set_state_consts for dt/2), the
            // centroid COM removal, back to beads, then onto the constraint surface
            launch_nm(c->p, c->pnm, nmm->F, nb, (size_t)3 * na, s);
            launch_pile(c->pnm, c->pile_c1, c->pile_c3, na, nb, 0, nz.take(draws), c->seed, c->step_ctr,
                        nz.stream_add++, c->com_parts, s);
            launch_com_remove(c->pnm, c->mass, c->com_parts, pile_blocks(na), 1.0 / c->mtot, na, s);
            launch_nm(c->pnm, c->p, nmm->B, nb, (size_t)3 * na, s);
            launch_rig_project(c->p, c->r, c->mass, na, nb, len2, tol, 0, c->rat_err, s);
        } else if (P.therm == RPMD_THERM_GLE) {
            // every channel onto the tangent space: the physical momenta (channel 0 is reloaded from
            // them by the next kick) and the auxiliary ones, so that the whole GLE process lives
            // on the constraint surface (projecting p alone leaves the auxiliary momenta feeding
            // energy into the constrained directions)
            launch_gle(c->p, c->gle_gp, c->gle_TS, c->mass, na, nb, c->gle_n1, nz.take(draws * c->gle_n1), c->seed,
                       c->step_ctr, nz.stream_add++, c->gle_parts, c->gle_pcom, 1.0 / sqrt(nb * c->mtot), s);
            launch_rig_project(c->p, c->r, c->mass, na, nb, len2, tol, 0, c->rat_err, s);
            for (int ch = 1; ch < c->gle_n1; ch++)
                launch_rig_project(c->gle_gp + (size_t)ch * draws, c->r, c->mass, na, nb, len2, tol, 1, c->rat_err, s);
        }
    };
    thermo();
    launch_kick(c->p, c->dvdr, halfdt, nullptr, 0.0, N, s);
    // nb <= 256 (every context): the sub-steps' blocks hold all beads of their molecules
    launch_rig_substeps(c->p, c->r, c->mass, na, nb, nb, halfsmalldt, smalldt, omegan, len2, tol, c->rat_err, s);
    RETCHK(forces_dev(c, c->r, nb, box, 1, c->dvdr, c->ev_step, check));
    launch_rig_final(c->p, c->r, c->dvdr, c->mass, na, nb, halfdt, dt, len2, tol, c->rat_err, c->rig_vparts,
                     c->ev_step + 50, s);
    thermo();
    launch_zero(c->dvdr2, N, s);
    launch_counter_inc(c->step_ctr, s);
    if (nz.stream_add > RPMD_STREAMS_PER_STEP) return RPMD_ERR_ARG;
    if (hipGetLastError() != hipSuccess) return RPMD_ERR_HIP;
    return 0;
}

// energies of the last rigid step: v = v_lf = forces(iopt = 1) energy, v_hf = 0,
// vir_lf, and vir_lf + vir_hf (src/evolve_rigid.f90:203-229)
int step_rig_energies(rpmd_ctx* c, double* v, double* vir_lfhf, double* vir_lf) {
    double h[60];
    HIPCHK(hipMemcpyAsync(h, c->ev_step, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double w[9];
    combine_ev(c, h, 1, c->st_box.L, v, w);
    for (int k = 0; k < 9; k++) {
        if (vir_lf) vir_lf[k] = w[k];
        if (vir_lfhf) vir_lfhf[k] = w[k] + h[50 + k];
    }
    return 0;
}

// a Verlet list that overflowed its capacity in any step since the lists were allocated
int vlist_status(rpmd_ctx* c) {
    if (c->oo_built) {
        int h = 0;
        HIPCHK(hipMemcpyAsync(&h, c->oo_ovf, sizeof h, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (h) return RPMD_ERR_NEIGH_OVERFLOW;
    }
    for (auto& x : c->vc) {
        if (!x.alloc) continue;
        int h[2];
        HIPCHK(hipMemcpyAsync(h, x.v.gate, sizeof h, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (h[1]) return RPMD_ERR_NEIGH_OVERFLOW;
    }
    return 0;
}

int rattle_status(rpmd_ctx* c) {
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->rat_err, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return h ? RPMD_ERR_RATTLE : 0;
}

// energies of the last step (src/evolve_pi.f90:166-169, 292-299)
int step_energies(rpmd_ctx* c, double* v, double* vew, double* vlj, double* vint, double* vir, double* vir_lf) {
    const rpmd_params& P = c->prm;
    const int mts = P.mts;
    std::vector<double> h((size_t)(mts + 4) * 50);
    HIPCHK(hipMemcpyAsync(h.data(), c->ev_step, h.size() * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const double* box = c->st_box.L;
    double vhf = 0, whf[9] = {0}, vi_last = 0;
    for (int im = 0; im < mts; im++) {
        double vv, w[9];
        combine_ev(c, &h[(size_t)im * 50], 4, box, &vv, w);
        for (int k = 0; k < 9; k++) whf[k] += w[k];
        vi_last = vv;
    }
    for (int k = 0; k < 9; k++) whf[k] /= (double)mts;
    const double* lf = &h[(size_t)mts * 50];
    double e_ew, w_ew[9], e_c, w_c[9], e_lj, w_lj[9], e_cb, w_cb[9];
    combine_ev(c, lf, 2, box, &e_ew, w_ew);
    combine_ev(c, lf + 50, 5, box, &e_c, w_c);
    combine_ev(c, c->lf_merged ? lf : lf + 100, 3, box, &e_lj, w_lj);
    combine_ev(c, lf + 150, 5, box, &e_cb, w_cb);
    double ew = e_ew, wew[9];
    for (int k = 0; k < 9; k++) wew[k] = w_ew[k];
    if (P.nbdf1 > 0 && P.sig > 0.0) {
        ew = ew - e_c + e_cb;
        for (int k = 0; k < 9; k++) wew[k] = wew[k] - w_c[k] + w_cb[k];
    }
    if (vew) *vew = ew;
    if (vlj) *vlj = e_lj;
    if (vint) *vint = vi_last;
    if (v) *v = ew + e_lj + vi_last;
    for (int k = 0; k < 9; k++) {
        double lf9 = w_lj[k] + wew[k];
        if (vir_lf) vir_lf[k] = lf9;
        if (vir) vir[k] = lf9 + whf[k];
    }
    return 0;
}

int set_state_consts(rpmd_ctx* c, const double* mass, double beta, double dt, int nb) {
    const rpmd_params& P = c->prm;
    const int na = c->na;
    c->hmass.assign(mass, mass + 3);
    c->mtot = 0.0;
    for (int j = 0; j < na; j++) c->mtot = c->mtot + mass[j];
    if (!get_nm(c, nb, 0)) return RPMD_ERR_NOMEM;
    std::vector<double> monod((size_t)12 * nb);
    ring_rpmd(mass, monod.data(), nb, dt / (double)P.mts, beta);
    HIPCHK(hipMemcpyAsync(c->monod, monod.data(), monod.size() * 8, hipMemcpyHostToDevice, c->st));
    std::vector<double> c1(nb), c3((size_t)3 * nb);
    // the kick's time step: dt / (2 mts) inside evolve_pi's MTS loop (src/evolve_pi.f90:103-106), dt / 2
    // at the two thermostat points of the rigid step
    pile_coeffs(P, mass, beta, nb, P.rigid ? 0.5 * dt : 0.5 * dt / (double)P.mts, c1.data(), c3.data());
    HIPCHK(hipMemcpyAsync(c->pile_c1, c1.data(), c1.size() * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pile_c3, c3.data(), c3.size() * 8, hipMemcpyHostToDevice, c->st));
    std::vector<double> c3m(na);
    // local Langevin: evolve_pi passes dt (src/evolve_pi.f90:75-77), evolve_rig_pi dt/2 (:126-130)
    const double gamma = 1.0 / (2.0 * P.ttau), betan = beta / (double)nb, dtl = P.rigid ? 0.5 * dt : dt;
    const double cc1 = exp(-0.5 * gamma * dtl), c2 = sqrt(1.0 - cc1 * cc1);
    for (int j = 0; j < na; j++) c3m[j] = c2 * sqrt(mass[j] / betan);
    HIPCHK(hipMemcpyAsync(c->lang_c3m, c3m.data(), c3m.size() * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->mass, mass, sizeof(double) * na, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return 0;
}

size_t noise_per_step(const rpmd_params& P, int na, int nb, int ngle) {
    const size_t d = (size_t)3 * na * nb;
    if (P.rigid) {
        if (P.therm == RPMD_THERM_PRG) return 2 * (d - (size_t)na - 2);  // ndof + 1 = d - 3 nm - 2
        if (P.therm == RPMD_THERM_PRL || P.therm == RPMD_THERM_PRO) return 2 * d;
        if (P.therm == RPMD_THERM_GLE) return 2 * d * (size_t)std::max(ngle, 1);
        return 0;
    }
    if (P.therm == RPMD_THERM_PRO) return 2 * (size_t)P.mts * d;
    if (P.therm == RPMD_THERM_PRL) return 2 * d;
    if (P.therm == RPMD_THERM_PRG) return 2 * (d - 2);
    if (P.therm == RPMD_THERM_PRM) return 2 * (size_t)P.mts * (d - 2);
    if (P.therm == RPMD_THERM_GLE) return 2 * d * (size_t)std::max(ngle, 1);
    return 0;
}

}  // namespace

// ---- allocation-free realft ---------------------------------------------------------------------
// For callers that cannot allocate device memory: the transform matrix and the staged rows live in
// this fixed device array (each GPU holds its own copy from module load to process end).  Host code
// never uses the array's name as an address: the device address comes from hipGetSymbolAddress, once
// per device, and is kept in noalloc_ptr[].
#define NOALLOC_MAT (256 * 256)     // the n x n matrix, n <= 256
#define NOALLOC_STAGE (1 << 19)     // doubles per staging half (rows x n of input, of output)
static __device__ __align__(256) double scratch[NOALLOC_MAT + 2 * NOALLOC_STAGE];

namespace {
#define NOALLOC_MAX_DEV 64
std::mutex noalloc_mu;
double* noalloc_ptr[NOALLOC_MAX_DEV] = {};
// the device address of scratch on the current device (looked up on its first call there)
int noalloc_scratch(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return RPMD_ERR_HIP;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RPMD_ERR_HIP;
        noalloc_ptr[dev] = (double*)p;
    }
    *out = noalloc_ptr[dev];
    return 0;
}
}  // namespace

// =================================================================================================
// C-ABI
// =================================================================================================
extern "C" {

const char* rpmd_strerror(int code) {
    switch (code) {
        case RPMD_OK: return "ok";
        case RPMD_ERR_NA_NOT_MULT3: return "potenl 1";
        case RPMD_ERR_NEIGH_OVERFLOW: return "neighbour list too small";
        case RPMD_ERR_NBDF_EVEN: return "* Beadiabatic factor must be ODD";
        case RPMD_ERR_NBDF_TOO_LARGE: return " * RP contraction beads must not exceed nb";
        case RPMD_ERR_RATTLE: return "max iterations exceeded in rattle";
        case RPMD_ERR_REALFT_N: return "realft 1";
        case RPMD_ERR_REALFT_MODE: return "realft 3";
        case RPMD_ERR_EWC_CUTOFF: return " ewc_basic : rcut too large ";
        case RPMD_ERR_SMEARED: return "ewc_correction not yet implemented for smeared charges";
        case RPMD_ERR_KMAX: return "Ewald kmax exceeds the device k tiles (31)";
        case RPMD_ERR_ARG: return "invalid argument";
        case RPMD_ERR_HIP: return "HIP runtime error";
        case RPMD_ERR_IOPT: return "invalid iopt";
        case RPMD_ERR_NOMEM: return "device allocation failed";
        case RPMD_ERR_THERM: return "thermostat not supported by this entry point";
        case RPMD_ERR_NOISE: return "noise buffer too short";
        case RPMD_ERR_GLE: return "GLE not initialised for this (na, nb): call rpmd_gle_init";
        case RPMD_ERR_TTAU: return "thermostat time constant ttau must be > 0";
        case RPMD_ERR_SOCKET: return "i-PI socket: no server at the address, or the connection broke";
        case RPMD_ERR_BARO: return " ** Invalid barostat **";
        case RPMD_ERR_IPI_NAT: return " @DRIVER MODE: Uh-oh! Particle number mismatch between i-pi and rpmd input!";
    }
    return "unknown error";
}

int rpmd_setup_ewald(int na, const double boxlxyz[3], rpmd_params* P) {
    if (!P || !boxlxyz || na <= 0) return RPMD_ERR_ARG;
    const double pi = acos(-1.0), cvar = 1.2;
    const double w = cvar * pow((double)na, -1.0 / 6.0);
    if (boxlxyz[0] == boxlxyz[1] && boxlxyz[0] == boxlxyz[2] && P->iamcub) {
        P->iamcub = 1;
        P->wrcut = std::min(0.5, w);
        if (P->ecut <= 0.0) {
            P->walpha = pi / P->wrcut;
            P->kmax = (int)(P->walpha);
            P->rkmax = 2.0 * pi * P->walpha;
        } else {
            double s = sqrt(-log(P->ecut));
            P->walpha = s / P->wrcut;
            P->kmax = (int)(s * P->walpha / pi);
            P->rkmax = 2.0 * s * P->walpha;
        }
    } else {
        P->iamcub = 0;
        const double rshort = std::min(boxlxyz[0], std::min(boxlxyz[1], boxlxyz[2]));
        const double rlong = std::max(boxlxyz[0], std::max(boxlxyz[1], boxlxyz[2]));
        P->wrcut = rshort * std::min(0.5, w);
        P->walpha = pi / P->wrcut;
        P->rkmax = 2.0 * pi * P->walpha;
        P->kmax = (int)(P->walpha * rlong);
    }
    return 0;
}

static void destroy_streams(rpmd_ctx* c) {
    for (hipEvent_t e : {c->ev_fork, c->ev_join, c->ev_fork3, c->ev_join3})
        if (e) hipEventDestroy(e);
    for (hipStream_t t : {c->st3, c->st2, c->st})
        if (t) hipStreamDestroy(t);
}

// every thermostat divides by ttau (PILE 2 pi / ttau, GLE 2 pi / ttau, Bussi dt / ttau, Langevin
// 1 / (2 ttau)): a zero or negative value would give infinite or NaN propagators
static bool ttau_ok(const rpmd_params* prm) { return prm->therm == RPMD_THERM_NONE || prm->ttau > 0.0; }

int rpmd_ctx_create(const rpmd_params* prm, int na, int nb_max, int precision, rpmd_ctx** out) {
    if (!prm || !out || na <= 0 || nb_max <= 0) return RPMD_ERR_ARG;
    if (!ttau_ok(prm)) return RPMD_ERR_TTAU;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (nb_max > 256) return RPMD_ERR_REALFT_N;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return RPMD_ERR_HIP;
    rpmd_ctx* c = new rpmd_ctx();
    c->prm = *prm;
    c->na = na;
    c->nb_max = nb_max;
    c->prec = precision;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st3, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork3, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join3, hipEventDisableTiming) != hipSuccess) {
        destroy_streams(c);
        delete c;
        return RPMD_ERR_HIP;
    }
    const size_t N = (size_t)3 * na * nb_max;
    const int nbl = nb_max;
    const int nck = kspace_nchunk(na);
    {
        // room for the cluster-sort grid of any box with aspect ratios up to 8 (at most 64^3)
        int p = 1;
        while (p < std::min(64, (int)ceil(8.0 * cbrt(std::max(na, 1) / 4.0)) + 1)) p <<= 1;
        c->ncell_cap = p * p * p;
    }
    c->off_real = 0;
    c->off_k = c->off_real + (size_t)nbl * ewald_real_parts(na) * 10;
    c->off_oo = c->off_k + (size_t)nbl * (RPMD_KMAX_DEV + 1) * 10;
    c->off_tem = c->off_oo + (size_t)nbl * oo_blocks(na) * 10;
    c->off_intra = c->off_tem + (size_t)nbl * tem_blocks(na) * 10;
    c->nparts = c->off_intra + (size_t)nbl * intra_blocks(na) * 10;
    int e = 0;
    e |= dalloc(&c->aos, N);
    e |= dalloc(&c->pos, N);
    e |= dalloc(&c->grad, N);
    e |= dalloc(&c->site, N);
    e |= dalloc(&c->gsite, N);
    e |= dalloc(&c->q, na);
    e |= dalloc(&c->mass, na);
    e |= dalloc(&c->cen, (size_t)3 * na);
    e |= dalloc(&c->parts, c->nparts);
    e |= dalloc(&c->ev, 50);
    e |= dalloc(&c->skp, (size_t)nbl * (RPMD_KMAX_DEV + 1) * nck * RPMD_KTILE * RPMD_KTILE * 2);
    e |= dalloc(&c->gk, (size_t)nbl * (RPMD_KMAX_DEV + 1) * RPMD_KTILE * RPMD_KTILE * 2);
    e |= dalloc(&c->gks, (size_t)nbl * (RPMD_KMAX_DEV + 1));
    e |= dalloc(&c->ph, (size_t)6 * na * nbl);
    e |= dalloc(&c->fpart, (size_t)(ewald_real_slots() + kspace_ngroups(RPMD_KMAX_DEV)) * 3 * na * nbl);
    e |= dalloc(&c->goo, (size_t)3 * (na / 3) * nbl);
    e |= dalloc(&c->kmask, (size_t)(RPMD_KMAX_DEV + 1) * RPMD_KTILE * RPMD_KTILE);
    e |= dalloc(&c->cell_of, (size_t)na * nbl);
    e |= dalloc(&c->cnt, (size_t)(c->ncell_cap + 1) * nbl);
    e |= dalloc(&c->start, (size_t)(c->ncell_cap + 1) * nbl);
    e |= dalloc(&c->perm, (size_t)na * nbl);
    e |= dalloc(&c->nnb, na);
    e |= dalloc(&c->nbl, (size_t)na * RPMD_NCAP);
    e |= dalloc(&c->ovf, 2);
    e |= dalloc(&c->oo_nnb, na / 3 + 1);
    e |= dalloc(&c->oo_nbl, (size_t)(na / 3 + 1) * RPMD_NCAP);
    e |= dalloc(&c->oo_ovf, 1);
    e |= dalloc(&c->oo_cb, (size_t)3 * (na / 3 + 1));
    e |= dalloc(&c->oo_dmax, 1);
    for (double** b : {&c->p, &c->r, &c->dvdr, &c->dvdr2, &c->dvdrl, &c->dvdre, &c->rst, &c->pnm, &c->rnm, &c->tmp,
                       &c->rb, &c->dvdrb, &c->tmp2})
        e |= dalloc(b, N);
    e |= dalloc(&c->monod, (size_t)12 * nbl);
    e |= dalloc(&c->pile_c1, nbl);
    e |= dalloc(&c->pile_c3, (size_t)3 * nbl);
    e |= dalloc(&c->com_parts, (size_t)3 * pile_blocks(na));
    e |= dalloc(&c->com_parts2, (size_t)3 * pile_blocks(na));
    e |= dalloc(&c->pstep_parts, (size_t)6 * pile_step_blocks(na));
    e |= dalloc(&c->lang_c3m, na);
    e |= dalloc(&c->step_ctr, 1);
    e |= dalloc(&c->bussi_parts, (size_t)bussi_parts_size());
    e |= dalloc(&c->bussi_alpha, 1);
    e |= dalloc(&c->gle_TS, 128);
    e |= dalloc(&c->gle_pcom, 24);
    e |= dalloc(&c->rat_err, 1);
    e |= dalloc(&c->rig_vparts, (size_t)7 * rigid_blocks(na) * nbl);
    e |= dalloc(&c->ev_step, (size_t)(64 + 4) * 50);
    if (e) {
        rpmd_ctx_destroy(c);
        return RPMD_ERR_NOMEM;
    }
    // initial contents, in the context stream's order (see vlist_for)
    launch_fill32(c->cnt, 0u, (size_t)(c->ncell_cap + 1) * nbl, c->st);
    launch_fill32(c->oo_ovf, 0u, 1, c->st);
    launch_fill32(c->oo_dmax, 0u, 2, c->st);
    launch_fill32(c->oo_cb, 0xffffffffu, (size_t)2 * 3 * (na / 3 + 1), c->st);
    launch_fill32(c->step_ctr, 0u, 2, c->st);
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        rpmd_ctx_destroy(c);
        return RPMD_ERR_HIP;
    }
    *out = c;
    return 0;
}

int rpmd_ctx_set_params(rpmd_ctx* c, const rpmd_params* prm) {
    if (!c || !prm) return RPMD_ERR_ARG;
    if (!ttau_ok(prm)) return RPMD_ERR_TTAU;
    c->prm = *prm;
    c->ks_valid = false;
    c->st_valid = false;
    invalidate(c);
    return 0;
}

int rpmd_ctx_set_precision(rpmd_ctx* c, int precision) {
    if (!c) return RPMD_ERR_ARG;
    c->prec = precision;
    invalidate(c);
    return 0;
}

int rpmd_ctx_destroy(rpmd_ctx* c) {
    if (!c) return 0;
    drop_graph(c);
    free_vcache(c);
    void* bufs[] = {c->ph, c->fpart, c->bussi_parts, c->bussi_alpha, c->gle_gp, c->gle_TS, c->baro_parts, c->baro_ev,
                    c->kmask_big, c->skp_big, c->gk_big, c->parts_big, c->fpart_big,
                    c->gle_parts, c->gle_pcom, c->rat_err, c->rig_vparts, c->aos, c->pos, c->grad, c->site, c->gsite, c->q, c->mass, c->cen, c->parts, c->ev,
                    c->skp, c->gk, c->gks, c->kmask, c->cell_of, c->cnt, c->start, c->perm, c->nnb, c->nbl, c->ovf,
                    c->oo_nnb, c->oo_nbl, c->oo_ovf, c->oo_cb, c->oo_dmax, c->p,
                    c->r, c->dvdr, c->dvdr2, c->dvdrl, c->dvdre, c->rst, c->pnm, c->rnm, c->tmp, c->rb, c->dvdrb,
                    c->tmp2, c->monod, c->pile_c1, c->pile_c3, c->com_parts, c->lang_c3m, c->noise,
                    c->step_ctr, c->ev_step, c->goo, c->com_parts2, c->pstep_parts};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (auto& t : c->ptab)
        if (t.d) hipFree(t.d);
    for (auto& kv : c->nm) {
        hipFree(kv.second.F);
        hipFree(kv.second.B);
        hipFree(kv.second.Fs);
        hipFree(kv.second.Bs);
        hipFree(kv.second.maps);
    }
    if (c->h_stage) hipHostFree(c->h_stage);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    destroy_streams(c);
    delete c;
    return 0;
}

int rpmd_forces(rpmd_ctx* c, const double* r, int nb, int na, const double boxlxyz[3], const double* z, int iopt,
                double* v, double* dvdr, double* virial) {
    if (!c || !r || !boxlxyz || !z || !v || !dvdr || !virial) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    if (iopt < 0 || iopt > 5) return RPMD_ERR_IOPT;
    // the call shares the lists, k set and charges with a device-resident state: graphs re-capture
    invalidate(c);
    RETCHK(upload_charges(c, z));
    RETCHK(upload_aos(c, r, c->pos, nb));
    RETCHK(forces_dev(c, c->pos, nb, boxlxyz, iopt, c->grad, c->ev, true));
    RETCHK(download_aos(c, c->grad, dvdr, nb));
    RETCHK(read_ev(c, c->ev, iopt, boxlxyz, v, virial));
    return 0;
}

int rpmd_full_forces(rpmd_ctx* c, const double* r, int na, int nb, double* v, double* vew, double* voo, double* vint,
                     double* vir, const double* z, const double boxlxyz[3], double* dvdr, double* dvdr2) {
    if (!c || !r || !z || !boxlxyz || !dvdr || !dvdr2) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    // a device-resident state keeps p, r, dvdr and dvdr2 between steps; this call works in buffers
    // the state does not own (the contraction's own scratch, rnm, tmp2, rb and dvdrb, is
    // rewritten by every step), so it never disturbs a state, but graphs are captured anew
    invalidate(c);
    double* gl = c->grad;
    RETCHK(upload_charges(c, z));
    RETCHK(upload_aos(c, r, c->pos, nb));
    RETCHK(full_forces_core(c, c->pos, nb, boxlxyz, &gl, c->pnm, c->dvdrl, c->dvdre, c->tmp, v, vew, voo, vint, vir));
    RETCHK(download_aos(c, gl, dvdr, nb));
    RETCHK(download_aos(c, c->pnm, dvdr2, nb));
    return 0;
}

int rpmd_realft(rpmd_ctx* c, double* data, int m, int n, int mode) {
    if (!c || !data || m <= 0 || n <= 0) return RPMD_ERR_ARG;
    if (n > 256) return RPMD_ERR_REALFT_N;
    if (mode != 1 && mode != -1) return RPMD_ERR_REALFT_MODE;
    const size_t N = (size_t)m * n;
    if (N > (size_t)3 * c->na * c->nb_max) return RPMD_ERR_ARG;
    NMMat* mm = get_nm(c, n, 0);
    if (!mm) return RPMD_ERR_NOMEM;
    RETCHK(ensure_stage(c, N));
    memcpy(c->h_stage, data, N * 8);
    HIPCHK(hipMemcpyAsync(c->tmp, c->h_stage, N * 8, hipMemcpyHostToDevice, c->st));
    launch_nm(c->tmp, c->tmp2, mode == 1 ? mm->F : mm->B, n, (size_t)m, c->st);
    HIPCHK(hipMemcpyAsync(c->h_stage, c->tmp2, N * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    memcpy(data, c->h_stage, N * 8);
    return 0;
}

int rpmd_realft_noalloc(double* data, int m, int n, int mode) {
    if (!data || m <= 0 || n <= 0) return RPMD_ERR_ARG;
    if (n > 256) return RPMD_ERR_REALFT_N;
    if (mode != 1 && mode != -1) return RPMD_ERR_REALFT_MODE;
    double* sp = nullptr;
    RETCHK(noalloc_scratch(&sp));
    double* M = sp;
    double* in = sp + NOALLOC_MAT;
    double* out = in + NOALLOC_STAGE;
    // the matrix, formed in static host storage (no heap), then the rows in chunks that fit the
    // staging halves: every copy and the kernel on the null stream, in order; the last copy waits
    static thread_local double hm[NOALLOC_MAT];
    realft_matrices_into(n, mode == 1 ? hm : nullptr, mode == 1 ? nullptr : hm);
    HIPCHK(hipMemcpy(M, hm, sizeof(double) * n * n, hipMemcpyHostToDevice));
    const int rows = NOALLOC_STAGE / n;
    for (int k0 = 0; k0 < m; k0 += rows) {
        const int cr = std::min(rows, m - k0);
        HIPCHK(hipMemcpy2D(in, sizeof(double) * cr, data + k0, sizeof(double) * m, sizeof(double) * cr, n,
                           hipMemcpyHostToDevice));
        launch_nm(in, out, M, n, (size_t)cr, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy2D(data + k0, sizeof(double) * m, out, sizeof(double) * cr, sizeof(double) * cr, n,
                           hipMemcpyDeviceToHost));
    }
    return 0;
}

int rpmd_free_ring(rpmd_ctx* c, double* p, double* r, const double* mass, int na, int nb, double dt, double beta) {
    if (!c || !p || !r || !mass || nb < 1 || beta <= 0.0) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb > c->nb_max) return RPMD_ERR_ARG;
    // scratch buffers only (pnm, rnm, and the monodromy in tmp2): a device-resident state is untouched
    std::vector<double> monod((size_t)12 * nb);
    ring_rpmd(mass, monod.data(), nb, dt, beta);
    HIPCHK(hipMemcpyAsync(c->tmp2, monod.data(), monod.size() * 8, hipMemcpyHostToDevice, c->st));
    RETCHK(upload_aos(c, p, c->pnm, nb));
    RETCHK(upload_aos(c, r, c->rnm, nb));
    launch_free_ring(c->pnm, c->rnm, c->tmp2, na, nb, c->st);
    RETCHK(download_aos(c, c->pnm, p, nb));
    RETCHK(download_aos(c, c->rnm, r, nb));
    return 0;
}

// one evolve_pi step on host arrays; with bc the barostat acts between the multiple time steps and
// the low-frequency forces and the box comes back in box_out
int evolve_pi_call(rpmd_ctx* c, double* p, double* r, double* v, double* vew, double* vlj, double* vint, double* dvdr,
                   double* dvdr2, double dt, const double* mass, int na, int nb, const double boxlxyz[3],
                   const double* z, double beta, double* vir, double* vir_lf, const double* noise, long long n_noise,
                   BaroCall* bc, double* box_out) {
    if (!c || !p || !r || !dvdr || !dvdr2 || !mass || !z || !boxlxyz) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    const size_t need = noise_per_step(c->prm, na, nb, c->gle_n1);
    if (noise && (size_t)n_noise < need) return RPMD_ERR_NOISE;
    RETCHK(upload_charges(c, z));
    // one host round trip: the uploaded arrays replace any device-resident state
    c->st_valid = false;
    invalidate(c);
    c->st_nb = nb;
    c->st_box = make_box(boxlxyz);
    c->st_beta = beta;
    c->st_dt = dt;
    RETCHK(set_state_consts(c, mass, beta, dt, nb));
    RETCHK(upload_aos(c, p, c->p, nb));
    RETCHK(upload_aos(c, r, c->r, nb));
    RETCHK(upload_aos(c, dvdr, c->dvdr, nb));
    RETCHK(upload_aos(c, dvdr2, c->dvdr2, nb));
    NoiseCursor nz{nullptr, 0, 0, 0};
    if (noise && need) {
        if (c->noise_cap < need) {
            if (c->noise) hipFree(c->noise);
            c->noise = nullptr;
            RETCHK(dalloc(&c->noise, need));
            c->noise_cap = need;
        }
        HIPCHK(hipMemcpyAsync(c->noise, noise, need * 8, hipMemcpyHostToDevice, c->st));
        nz.base = c->noise;
        nz.cap = need;
    }
    RETCHK(step_dev(c, nz, true, bc));
    RETCHK(download_aos(c, c->p, p, nb));
    RETCHK(download_aos(c, c->r, r, nb));
    RETCHK(download_aos(c, c->dvdr, dvdr, nb));
    RETCHK(download_aos(c, c->dvdr2, dvdr2, nb));
    RETCHK(step_energies(c, v, vew, vlj, vint, vir, vir_lf));
    if (box_out)
        for (int d = 0; d < 3; d++) box_out[d] = c->st_box.L[d];
    c->st_valid = false;
    return 0;
}

int rpmd_evolve_pi(rpmd_ctx* c, double* p, double* r, double* v, double* vew, double* vlj, double* vint, double* dvdr,
                   double* dvdr2, double dt, const double* mass, int na, int nb, const double boxlxyz[3],
                   const double* z, double beta, double* vir, double* vir_lf, const double* noise, long long n_noise) {
    return evolve_pi_call(c, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir, vir_lf,
                          noise, n_noise, nullptr, nullptr);
}

int rpmd_set_barostat(rpmd_ctx* c, int baro, double taub, double patm, double dv) {
    if (!c) return RPMD_ERR_ARG;
    if (baro != RPMD_BARO_NONE && baro != RPMD_BARO_BER && baro != RPMD_BARO_MCI) return RPMD_ERR_BARO;
    if (baro != RPMD_BARO_NONE && !(taub > 0.0)) return RPMD_ERR_ARG;
    c->baro = baro;
    c->taub = taub;
    c->patm = patm;
    c->dv = dv;
    c->nacc = c->nmove = 0;
    return 0;
}

int rpmd_barostat_counts(const rpmd_ctx* c, long long* nacc, long long* nmove) {
    if (!c) return RPMD_ERR_ARG;
    if (nacc) *nacc = c->nacc;
    if (nmove) *nmove = c->nmove;
    return 0;
}

int rpmd_evolve_pi_npt(rpmd_ctx* c, double* p, double* r, double* v, double* vew, double* vlj, double* vint,
                       double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb, double boxlxyz[3],
                       const double* z, double beta, double* vir, double* vir_lf, const double* noise,
                       long long n_noise, const double* uniforms, int n_uniforms) {
    if (!c || !vir_lf || !boxlxyz) return RPMD_ERR_ARG;
    if (c->baro == RPMD_BARO_NONE) return RPMD_ERR_BARO;
    if (c->prm.rigid || (uniforms && n_uniforms < 0)) return RPMD_ERR_ARG;
    BaroCall bc;
    memcpy(bc.vir_lf, vir_lf, sizeof bc.vir_lf);
    bc.u = uniforms;
    bc.nu = uniforms ? n_uniforms : 0;
    if (c->baro == RPMD_BARO_BER) RETCHK(baro_buffers(c, std::max(c->prm.mts, 1)));
    double box[3] = {boxlxyz[0], boxlxyz[1], boxlxyz[2]};
    RETCHK(evolve_pi_call(c, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, noise,
                          n_noise, &bc, boxlxyz));
    return 0;
}

// one rigid step on host arrays; with bc the barostat acts after it (src/evolve_rigid.f90:203-220)
// and the box comes back in box_out
int evolve_rig_pi_call(rpmd_ctx* c, double* p, double* r, double* v, double* v_lf, double* v_hf, double* dvdr,
                       double* dvdr2, double dt, const double* mass, int na, int nb, const double boxlxyz[3],
                       const double* z, double beta, double* vir, double* vir_lf, const double* noise,
                       long long n_noise, BaroCall* bc, double* box_out) {
    if (!c || !p || !r || !dvdr || !dvdr2 || !mass || !z || !boxlxyz || !vir) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    if (!c->prm.rigid) return RPMD_ERR_ARG;
    const size_t need = noise_per_step(c->prm, na, nb, c->gle_n1);
    if (noise && (size_t)n_noise < need) return RPMD_ERR_NOISE;
    RETCHK(upload_charges(c, z));
    // one host round trip: the uploaded arrays replace any device-resident state
    c->st_valid = false;
    invalidate(c);
    c->st_nb = nb;
    c->st_box = make_box(boxlxyz);
    c->st_beta = beta;
    c->st_dt = dt;
    RETCHK(set_state_consts(c, mass, beta, dt, nb));
    RETCHK(upload_aos(c, p, c->p, nb));
    RETCHK(upload_aos(c, r, c->r, nb));
    RETCHK(upload_aos(c, dvdr, c->dvdr, nb));
    launch_zero_int(c->rat_err, 1, c->st);
    NoiseCursor nz{nullptr, 0, 0, 0};
    if (noise && need) {
        if (c->noise_cap < need) {
            if (c->noise) hipFree(c->noise);
            c->noise = nullptr;
            RETCHK(dalloc(&c->noise, need));
            c->noise_cap = need;
        }
        HIPCHK(hipMemcpyAsync(c->noise, noise, need * 8, hipMemcpyHostToDevice, c->st));
        nz.base = c->noise;
        nz.cap = need;
    }
    RETCHK(step_rig_dev(c, nz, true));
    RETCHK(rattle_status(c));
    double vv, w[9], wl[9];
    RETCHK(step_rig_energies(c, &vv, w, wl));
    for (int k = 0; k < 9; k++) {
        vir[k] = w[k] + vir[k];
        if (vir_lf) vir_lf[k] = wl[k];
    }
    if (bc) RETCHK(rig_baro_tail(c, bc, nb, &vv, vir));
    RETCHK(download_aos(c, c->p, p, nb));
    RETCHK(download_aos(c, c->r, r, nb));
    RETCHK(download_aos(c, c->dvdr, dvdr, nb));
    RETCHK(download_aos(c, c->dvdr2, dvdr2, nb));
    if (v) *v = vv;
    if (v_lf) *v_lf = vv;
    if (v_hf) *v_hf = 0.0;
    if (box_out)
        for (int d = 0; d < 3; d++) box_out[d] = c->st_box.L[d];
    c->st_valid = false;
    return 0;
}

int rpmd_evolve_rig_pi(rpmd_ctx* c, double* p, double* r, double* v, double* v_lf, double* v_hf, double* dvdr,
                       double* dvdr2, double dt, const double* mass, int na, int nb, const double boxlxyz[3],
                       const double* z, double beta, double* vir, double* vir_lf, const double* noise,
                       long long n_noise) {
    return evolve_rig_pi_call(c, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir, vir_lf,
                              noise, n_noise, nullptr, nullptr);
}

int rpmd_evolve_rig_pi_npt(rpmd_ctx* c, double* p, double* r, double* v, double* v_lf, double* v_hf, double* dvdr,
                           double* dvdr2, double dt, const double* mass, int na, int nb, double boxlxyz[3],
                           const double* z, double beta, double* vir, double* vir_lf, const double* noise,
                           long long n_noise, const double* uniforms, int n_uniforms) {
    if (!c || !boxlxyz || !vir) return RPMD_ERR_ARG;
    if (c->baro == RPMD_BARO_NONE) return RPMD_ERR_BARO;
    if (uniforms && n_uniforms < 0) return RPMD_ERR_ARG;
    BaroCall bc;
    bc.u = uniforms;
    bc.nu = uniforms ? n_uniforms : 0;
    if (c->baro == RPMD_BARO_BER) RETCHK(baro_buffers(c, 1));
    double box[3] = {boxlxyz[0], boxlxyz[1], boxlxyz[2]};
    return evolve_rig_pi_call(c, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, noise,
                              n_noise, &bc, boxlxyz);
}


int rpmd_state_upload(rpmd_ctx* c, const double* p, const double* r, const double* mass, const double* z, int na,
                      int nb, const double boxlxyz[3], double beta, double dt, unsigned long long seed) {
    if (!c || !p || !r || !mass || !z || !boxlxyz) return RPMD_ERR_ARG;
    if (na % 3) return RPMD_ERR_NA_NOT_MULT3;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    invalidate(c);
    RETCHK(upload_charges(c, z));
    c->st_z.assign(z, z + na);
    c->st_nb = nb;
    c->st_box = make_box(boxlxyz);
    c->st_beta = beta;
    c->st_dt = dt;
    c->seed = seed;
    RETCHK(set_state_consts(c, mass, beta, dt, nb));
    RETCHK(upload_aos(c, p, c->p, nb));
    RETCHK(upload_aos(c, r, c->r, nb));
    launch_fill32(c->step_ctr, 0u, 2, c->st);
    launch_zero_int(c->rat_err, 1, c->st);
    // initial forces: full_forces semantics for the uncontracted case (src/md_forces.f90:305-472)
    const size_t N = (size_t)3 * na * nb;
    RETCHK(forces_dev(c, c->r, nb, boxlxyz, 4, c->dvdr2, c->ev, true));
    RETCHK(forces_dev(c, c->r, nb, boxlxyz, 2, c->dvdr, c->ev, true));
    RETCHK(forces_dev(c, c->r, nb, boxlxyz, 3, c->dvdrl, c->ev, true));
    launch_axpy(c->dvdr, c->dvdrl, 1.0, N, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    c->st_valid = true;
    return 0;
}

int rpmd_sample_rpmd(rpmd_ctx* c, double* p, int na, int nb, const double* mass, double beta, const double* noise,
                     long long n_noise, unsigned long long seed) {
    if (!c || !p || !mass || beta <= 0.0) return RPMD_ERR_ARG;
    if (na != c->na || nb < 1 || nb > c->nb_max) return RPMD_ERR_ARG;
    const size_t N = (size_t)3 * na * nb;
    if (noise && (size_t)n_noise < N) return RPMD_ERR_NOISE;
    const double* dn = nullptr;
    if (noise) {
        if (c->noise_cap < N) {
            if (c->noise) hipFree(c->noise);
            c->noise = nullptr;
            RETCHK(dalloc(&c->noise, N) ? RPMD_ERR_NOMEM : 0);
            c->noise_cap = N;
        }
        HIPCHK(hipMemcpyAsync(c->noise, noise, N * 8, hipMemcpyHostToDevice, c->st));
        dn = c->noise;
    }
    double totmass = 0.0;
    for (int j = 0; j < na; j++) totmass = totmass + mass[j];
    totmass = (double)nb * totmass;
    // scratch only (a device-resident state keeps its own masses): the masses in rnm, p in tmp2,
    // the per-block partial sums (3 per block) and the shift velocity in tmp
    HIPCHK(hipMemcpyAsync(c->rnm, mass, sizeof(double) * na, hipMemcpyHostToDevice, c->st));
    const size_t nparts = (size_t)3 * pile_blocks(na) * nb;
    launch_sample_rpmd(c->tmp2, c->rnm, na, nb, beta, totmass, dn, seed, ~1ull, c->tmp, c->tmp + nparts, c->st);
    RETCHK(download_aos(c, c->tmp2, p, nb));
    return 0;
}

int rpmd_gle_init(rpmd_ctx* c, int ns, const double* A, int na, int nb, double dt, double beta, const double* gp,
                  unsigned long long seed) {
    if (!c || !A || ns < 0 || ns > 7 || nb < 1 || nb > c->nb_max || beta <= 0.0) return RPMD_ERR_ARG;
    if (na != c->na) return RPMD_ERR_ARG;
    if (!(c->prm.ttau > 0.0)) return RPMD_ERR_TTAU;  // A is scaled by 2 pi / ttau
    const int n1 = ns + 1;
    const size_t N = (size_t)3 * na * nb;
    double TS[128] = {0};
    gle_propagators(A, n1, c->prm.ttau, nb, dt, beta, TS, TS + 64);
    if (c->gle_gp && c->gle_nb * (size_t)c->gle_n1 < (size_t)nb * n1) {
        hipFree(c->gle_gp);
        hipFree(c->gle_parts);
        c->gle_gp = c->gle_parts = nullptr;
    }
    if (!c->gle_gp) {
        RETCHK(dalloc(&c->gle_gp, N * n1) ? RPMD_ERR_NOMEM : 0);
        RETCHK(dalloc(&c->gle_parts, (size_t)gle_parts_size(na, nb)) ? RPMD_ERR_NOMEM : 0);
    }
    HIPCHK(hipMemcpyAsync(c->gle_TS, TS, sizeof TS, hipMemcpyHostToDevice, c->st));
    if (gp) {
        // reference layout gp(3,na,nb,ns+1) -> channel-major SoA per bead
        std::vector<double> h(N * n1);
        for (int i = 0; i < n1; i++)
            for (int b = 0; b < nb; b++)
                for (int a = 0; a < na; a++)
                    for (int d = 0; d < 3; d++)
                        h[(size_t)i * N + ((size_t)b * 3 + d) * na + a] = gp[(size_t)i * N + ((size_t)b * na + a) * 3 + d];
        HIPCHK(hipMemcpyAsync(c->gle_gp, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    } else {
        // C = (nb/beta) I, so chol(C) xi = sqrt(nb/beta) xi; a stream id apart from the step streams
        launch_gle_draw(c->gle_gp, sqrt((double)nb / beta), na, nb, n1, seed, ~0ull, c->st);
        HIPCHK(hipStreamSynchronize(c->st));
    }
    c->gle_n1 = n1;
    c->gle_nb = nb;
    invalidate(c);
    return 0;
}

int rpmd_ctx_set_streams(rpmd_ctx* c, int streams) {
    if (!c || streams < 1 || streams > 3) return RPMD_ERR_ARG;
    c->streams = streams;
    drop_graph(c);  // replica batches run one stream per replica and keep theirs
    return 0;
}

int rpmd_use_graph(rpmd_ctx* c, int enable) {
    if (!c) return RPMD_ERR_ARG;
    c->use_graph = enable;
    if (!enable) drop_graph(c);
    return 0;
}

namespace {
int run_status(rpmd_ctx* c) {
    RETCHK(vlist_status(c));
    return c->prm.rigid ? rattle_status(c) : 0;
}
int step_any(rpmd_ctx* c, NoiseCursor& nz) {
    return c->prm.rigid ? step_rig_dev(c, nz, false) : step_dev(c, nz, false, nullptr);
}
}  // namespace

int rpmd_run_steps(rpmd_ctx* c, int nsteps) {
    if (!c || !c->st_valid || nsteps < 0) return RPMD_ERR_ARG;
    // the state's own charges (a host call may have uploaded others; then the epoch moved too)
    RETCHK(upload_charges(c, c->st_z.data()));
    auto eager = [&](int k) -> int {
        for (int i = 0; i < k; i++) {
            NoiseCursor nz{nullptr, 0, 0, 0};
            RETCHK(step_any(c, nz));
        }
        return 0;
    };
    int left = nsteps;
    if (c->use_graph && !c->prof && left >= 2) {
        if (c->gexec && c->g_epoch != c->epoch) drop_graph(c);
        if (!c->gexec) {
            // one eager step first: it brings every list, table and key up to date outside the
            // capture, so the captured steps hold no forced rebuild; then two steps are captured
            // (each step exchanges r and rst, so two return the buffers to their roles)
            RETCHK(eager(1));
            left--;
            if (left >= 2) {
                HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                c->g_roles = roles_of(c);
                int rc = 0;
                for (int i = 0; i < 2 && !rc; i++) {
                    NoiseCursor nz{nullptr, 0, 0, 0};
                    rc = step_any(c, nz);
                }
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(c->st, &g);
                if (rc || e != hipSuccess) {
                    if (g) hipGraphDestroy(g);
                    // capture leaves the host-side roles as after two steps: restore them
                    c->r = c->g_roles.a[0], c->rst = c->g_roles.a[1], c->dvdr = c->g_roles.a[2];
                    c->dvdrl = c->g_roles.a[3], c->tmp = c->g_roles.a[4];
                    return rc ? rc : RPMD_ERR_HIP;
                }
                c->graph = g;
                if (hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0) != hipSuccess) {
                    drop_graph(c);
                    return RPMD_ERR_HIP;
                }
                c->g_epoch = c->epoch;
                // capturing ran no kernel but exchanged the roles twice: they are as captured
            }
        }
        if (c->gexec && left >= 2 && roles_of(c) != c->g_roles) {
            // an odd number of steps ran since the capture: one eager step brings the roles back
            RETCHK(eager(1));
            left--;
            if (roles_of(c) != c->g_roles) drop_graph(c);
        }
        if (c->gexec && roles_of(c) == c->g_roles)
            for (; left >= 2; left -= 2) HIPCHK(hipGraphLaunch(c->gexec, c->st));
    }
    RETCHK(eager(left));
    HIPCHK(hipStreamSynchronize(c->st));
    return run_status(c);
}

// ---- replica batches ---------------------------------------------------------------------------
// Independent trajectories (each its own context: box, Ewald set-up, na, nb, temperature) stepped
// together: one hipGraph per batch holds two steps of every replica, each replica on its own
// stream branch (fork from the batch stream, join back), so one graph launch advances all of them
// and the replicas' kernels overlap on the GPU.
// the state of one replica between the two halves of a batched step
struct FusedRep {
    ReduceJobs red;
    int use_list = 0;
    unsigned long long* oo_dmax = nullptr;
};

struct rpmd_batch {
    std::vector<rpmd_ctx*> ctx;
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // per replica at capture: the context's epoch and buffer roles
    std::vector<uint64_t> epoch;
    std::vector<Roles> roles;
    // batched steps (every replica takes the fused PILE-L step in mixed precision): the four force
    // kernels of all replicas as one launch each, on st (and st2 for the pair walk beside the k-space
    // chain, if the autotuner chose that); ev_a per replica: its first half done
    bool fused = false;
    ForceBatch* fb = nullptr;
    hipStream_t st2 = nullptr;
    hipEvent_t ev_b = nullptr, ev_f2 = nullptr, ev_j2 = nullptr;
    std::vector<hipEvent_t> ev_a;
    std::vector<FusedRep> fr;
    std::vector<FBItem> items;
    // the autotuner's choice: cap of the pair walk (blocks per CU, 0 none), pair walk on st2
    bool tuned = false;
    int bpc = 0, two = 0;
    // the batch autotuner's choice between the batched step and each replica's own step (the
    // same bits either way), made once by timing graph replays of both
    bool mode_tuned = false, mode_fused = true;
    float ms_fused = 0.0f, ms_separate = 0.0f;
};

// a replica that the batched step takes: mixed precision, flexible molecules, PILE-L with one time
// step and no contraction (the fused PILE step, up to 64 beads), cubic Ewald without smeared charges
static bool fused_ok(const rpmd_ctx* c) {
    const rpmd_params& P = c->prm;
    return c->prec == RPMD_PREC_MIXED && !P.rigid && P.therm == RPMD_THERM_PRO && P.mts == 1 && P.nbdf1 <= 0 &&
           P.nbdf2 <= 0 && P.iamcub && c->st_nb >= 1 && c->st_nb <= 64 && !(P.wm > 0.0 || P.wh > 0.0) && !c->prof &&
           P.kmax >= 0 && P.kmax <= RPMD_KMAX_DEV;
}

// The first half of the replica's step on its stream, as step_dev's fused PILE step with one stream:
// the PILE step, the intramolecular forces, then forces(iopt = 1) up to its four dominant kernels,
// whose inputs go to it (forces_dev's single-stream sequence: O-O centroid list, k set, pair
// table, real-space list, sites and phases, list upkeep).
static int fused_first_half(rpmd_ctx* c, FusedRep& fr, FBItem& it) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nb = c->st_nb;
    hipStream_t s = c->st;
    const double dt = c->st_dt, halfdt = 0.5 * dt, halfdtsmall = 0.5 * dt;  // mts = 1
    const double* box = c->st_box.L;
    NMMat* nmm = get_nm(c, nb, 0);
    if (!nmm) return RPMD_ERR_NOMEM;
    NoiseCursor nz{nullptr, 0, 0, 0};
    const size_t draws = (size_t)3 * na * nb;
    const double* n1 = nz.take(draws);
    const double* n2 = nz.take(draws);
    const uint64_t sa1 = nz.stream_add++;
    (void)nz.stream_add++;
    if (!launch_pile_step(c->p, c->dvdr, halfdt, c->dvdr2, halfdtsmall, c->r, c->p, c->rst, nmm->F, nmm->B, nb, na,
                          c->monod, c->mass, 1.0 / c->mtot, c->pile_c1, c->pile_c3, n1, n2, c->seed, c->step_ctr, sa1,
                          c->pstep_parts, s))
        return RPMD_ERR_ARG;
    fr = FusedRep{};
    RETCHK(forces_dev(c, c->rst, nb, box, 4, c->dvdr2, c->ev_step, false, &fr.red, nullptr));
    std::swap(c->r, c->rst);
    c->lf_merged = true;
    const double* pos = c->r;
    const Box bx = make_box(box);
    const double boxmin = std::min(box[0], std::min(box[1], box[2]));
    const double oo_cut = std::min(P.rcut, 0.5 * boxmin);
    if (nb > 3) {
        const double skin = 1.0 / 0.5291772108;
        const double key[5] = {box[0], box[1], box[2], oo_cut, skin};
        const int force = (!c->oo_built || memcmp(key, c->oo_key, sizeof key) != 0) ? 1 : 0;
        memcpy(c->oo_key, key, sizeof key);
        c->oo_built = true;
        launch_centroid_vl(pos, c->cen, na, nb, c->oo_cb, c->oo_dmax, s);
        launch_nlist_vl(c->cen, na, bx, oo_cut, skin, force, c->oo_nnb, c->oo_nbl, c->oo_cb, c->oo_dmax, c->oo_ovf,
                        c->nnb, c->nbl, s);
        fr.use_list = 1;
        fr.oo_dmax = c->oo_dmax;
    }
    RETCHK(build_kset(c, box));
    const Box sb = unit_box();
    const PairTab* ptab = nullptr;
    RETCHK(pair_table_for(c, sb, P.wrcut, P.walpha, s, &ptab));
    CellGrid g = make_grid(sb, na);
    if (morton_bins(g) > c->ncell_cap) return RPMD_ERR_ARG;
    const double skin = (1.0 / box[0]) * (EW_SKIN_A / 0.5291772108);
    VList* vl = nullptr;
    int force = 0;
    RETCHK(vlist_for(c, nb, box, P.wrcut, skin, sb, &vl, &force, 0));
    launch_sites_vl(pos, c->site, na, nb, P.alpha, 1, bx, c->ks.lat, c->ph, c->q, *vl, !force, sb, s);
    launch_vlist(*vl, c->site, c->q, na, nb, sb, g, P.wrcut, skin, force, c->cell_of, c->cnt, c->start, s);
    it = FBItem{vl, na, nb, sb, P.wrcut, P.walpha, c->fpart, c->parts + c->off_real, ptab, c->ph, c->q, &c->ks,
                c->skp, c->kmask, c->gk, c->gks, c->parts + c->off_k,
                c->fpart + (size_t)ewald_real_slots() * 3 * na * nb};
    return hipGetLastError() == hipSuccess ? 0 : RPMD_ERR_HIP;
}

// The second half: the O-O pairs with the chain rule of the Ewald slots, the energy jobs, the closing
// kicks and the reductions, as the single-stream step after its force kernels.
static int fused_second_half(rpmd_ctx* c, FusedRep& fr) {
    const rpmd_params& P = c->prm;
    const int na = c->na, nb = c->st_nb;
    hipStream_t s = c->st;
    const double dt = c->st_dt, halfdt = 0.5 * dt, halfdtsmall = 0.5 * dt;
    const double* box = c->st_box.L;
    const size_t N = (size_t)3 * na * nb;
    const Box bx = make_box(box);
    const double sc = 1.0 / box[0];
    const int chain_npart = ewald_real_slots() + 1;
    ReduceJobs jobs;
    jobs.add(c->parts + c->off_real, nb * ewald_real_parts(na), 0, sc / nb);
    jobs.add(c->parts + c->off_k, nb * (P.kmax + 1), 10, sc / nb);
    double* po = c->parts + c->off_oo;
    launch_oo<float>(c->r, na, nb, bx, c->nnb, c->nbl, fr.use_list, P.rcut, P, c->dvdr, po, s, c->fpart, chain_npart,
                     P.alpha, sc * sc, 0, fr.oo_dmax);
    jobs.add(po, nb * oo_blocks(na), 20, 1.0 / nb);
    const int base = 50;  // the forces(iopt = 1) slot follows the one of the intramolecular forces
    for (int j = 0; j < jobs.n; j++) {
        if (fr.red.n >= RED_JOBS) return RPMD_ERR_ARG;
        fr.red.add(jobs.parts[j], jobs.nparts[j], base + jobs.out[j], jobs.scale[j]);
    }
    launch_kick(c->p, c->dvdr2, halfdtsmall, c->dvdr, halfdt, N, s, c->step_ctr);
    launch_reduce_ev(fr.red, c->ev_step, s);
    return hipGetLastError() == hipSuccess ? 0 : RPMD_ERR_HIP;
}

// The autotuner of the batched force launches: each choice of the pair walk's cap (none, 6 or 4
// blocks per CU) and stream layout (after the k-space chain, or beside it on a second stream) is
// launched three times and timed with events on the batch stream; the fastest is kept.  The
// kernels only overwrite their outputs, so the extra launches leave the step's results unchanged.
static int fused_tune(rpmd_batch* b) {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return RPMD_ERR_HIP;
    float best = 1e30f;
    int rc = 0;
    for (int two = 0; two < 2 && !rc; two++)
        for (int bpc : {0, 6, 4}) {
            auto once = [&]() {
                if (two) {
                    (void)hipEventRecord(b->ev_f2, b->st);
                    (void)hipStreamWaitEvent(b->st2, b->ev_f2, 0);
                    force_batch_launch(b->fb, b->st, b->st2, bpc);
                    (void)hipEventRecord(b->ev_j2, b->st2);
                    (void)hipStreamWaitEvent(b->st, b->ev_j2, 0);
                } else {
                    force_batch_launch(b->fb, b->st, b->st, bpc);
                }
            };
            once();  // warm
            if (hipEventRecord(t0, b->st) != hipSuccess) rc = RPMD_ERR_HIP;
            for (int k = 0; k < 3; k++) once();
            float ms = 0.0f;
            if (rc || hipEventRecord(t1, b->st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess) {
                rc = RPMD_ERR_HIP;
                break;
            }
            if (ms < best) best = ms, b->bpc = bpc, b->two = two;
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    b->tuned = !rc;
    return rc;
}

// one batched step of every replica: first halves on the replicas' streams, the four force kernels
// of all replicas as four launches, second halves on the replicas' streams, all joined on b->st
static int fused_step(rpmd_batch* b) {
    const int n = (int)b->ctx.size();
    b->fr.resize(n);
    b->items.resize(n);
    HIPCHK(hipEventRecord(b->ev[2 * n], b->st));
    for (int i = 0; i < n; i++) {
        rpmd_ctx* c = b->ctx[i];
        HIPCHK(hipStreamWaitEvent(c->st, b->ev[2 * n], 0));
        RETCHK(fused_first_half(c, b->fr[i], b->items[i]));
        HIPCHK(hipEventRecord(b->ev_a[i], c->st));
        HIPCHK(hipStreamWaitEvent(b->st, b->ev_a[i], 0));
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(b->st, &cs));
    const int rc = force_batch_set(b->fb, b->items.data(), n, b->st, cs != hipStreamCaptureStatusNone);
    if (rc < 0) return RPMD_ERR_HIP;
    if (rc > 0) return RPMD_ERR_ARG;  // the records changed under capture: the caller steps eagerly
    if (!b->tuned && cs == hipStreamCaptureStatusNone) RETCHK(fused_tune(b));
    if (b->two) {
        HIPCHK(hipEventRecord(b->ev_f2, b->st));
        HIPCHK(hipStreamWaitEvent(b->st2, b->ev_f2, 0));
        force_batch_launch(b->fb, b->st, b->st2, b->bpc);
        HIPCHK(hipEventRecord(b->ev_j2, b->st2));
        HIPCHK(hipStreamWaitEvent(b->st, b->ev_j2, 0));
    } else {
        force_batch_launch(b->fb, b->st, b->st, b->bpc);
    }
    HIPCHK(hipEventRecord(b->ev_b, b->st));
    for (int i = 0; i < n; i++) {
        rpmd_ctx* c = b->ctx[i];
        HIPCHK(hipStreamWaitEvent(c->st, b->ev_b, 0));
        RETCHK(fused_second_half(c, b->fr[i]));
        HIPCHK(hipEventRecord(b->ev[2 * i], c->st));
        HIPCHK(hipStreamWaitEvent(b->st, b->ev[2 * i], 0));
    }
    return 0;
}

int rpmd_batch_create(rpmd_ctx* const* ctxs, int n, rpmd_batch** out) {
    if (!ctxs || n < 1 || !out) return RPMD_ERR_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !ctxs[i]->st_valid) return RPMD_ERR_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return RPMD_ERR_ARG;
    }
    rpmd_batch* b = new rpmd_batch();
    b->ctx.assign(ctxs, ctxs + n);
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking) != hipSuccess) {
        rpmd_batch_destroy(b);
        return RPMD_ERR_HIP;
    }
    b->ev.resize((size_t)2 * n + 1, nullptr);
    b->ev_a.resize(n, nullptr);
    for (auto* v : {&b->ev, &b->ev_a})
        for (auto& e : *v)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                rpmd_batch_destroy(b);
                return RPMD_ERR_HIP;
            }
    for (hipEvent_t* e : {&b->ev_b, &b->ev_f2, &b->ev_j2})
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
            rpmd_batch_destroy(b);
            return RPMD_ERR_HIP;
        }
    b->fb = force_batch_new();
    *out = b;
    return 0;
}

static void batch_drop_graph(rpmd_batch* b) {
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

int rpmd_batch_destroy(rpmd_batch* b) {
    if (!b) return 0;
    batch_drop_graph(b);
    for (auto* v : {&b->ev, &b->ev_a})
        for (hipEvent_t e : *v)
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->ev_b, b->ev_f2, b->ev_j2})
        if (e) (void)hipEventDestroy(e);
    if (b->st) (void)hipStreamDestroy(b->st);
    if (b->st2) (void)hipStreamDestroy(b->st2);
    force_batch_free(b->fb);
    delete b;
    return 0;
}

int rpmd_batch_run_steps(rpmd_batch* b, int nsteps) {
    if (!b || nsteps < 0) return RPMD_ERR_ARG;
    const int n = (int)b->ctx.size();
    for (rpmd_ctx* c : b->ctx) {
        if (!c->st_valid) return RPMD_ERR_ARG;
        RETCHK(upload_charges(c, c->st_z.data()));
    }
    // each replica runs on its own context stream only: the replicas already run side by side, and
    // a second stream per replica made the batch 1.8x slower (48 replicas of config 3).
    // k[i] steps of replica i, forked from and joined back to the batch stream.
    auto fork_steps = [&](const std::vector<int>& k) -> int {
        HIPCHK(hipEventRecord(b->ev[2 * n], b->st));
        for (int i = 0; i < n; i++) {
            if (k[i] <= 0) continue;
            rpmd_ctx* c = b->ctx[i];
            HIPCHK(hipStreamWaitEvent(c->st, b->ev[2 * n], 0));
            const int streams = c->streams;
            c->streams = 1;
            int rc = 0;
            for (int q = 0; q < k[i] && !rc; q++) {
                NoiseCursor nz{nullptr, 0, 0, 0};
                rc = step_any(c, nz);
            }
            c->streams = streams;
            RETCHK(rc);
            HIPCHK(hipEventRecord(b->ev[2 * i], c->st));
            HIPCHK(hipStreamWaitEvent(b->st, b->ev[2 * i], 0));
        }
        return 0;
    };
    std::vector<int> left(n, nsteps), one(n, 1);
    // every replica on the batched step (one launch per force kernel for all of them), when each
    // takes it and the autotuner has not found the replicas' own steps side by side faster
    bool eligible = true;
    for (rpmd_ctx* c : b->ctx) eligible = eligible && fused_ok(c);
    const bool fused = eligible && (!b->mode_tuned || b->mode_fused);
    auto steps_in = [&](bool fz, int k) -> int {
        if (!fz) return fork_steps(std::vector<int>(n, k));
        for (int q = 0; q < k; q++) RETCHK(fused_step(b));
        return 0;
    };
    // two steps of every replica captured into a graph; the buffer roles and epochs at capture
    auto capture = [&](bool fz, hipGraph_t* gr, hipGraphExec_t* ge) -> int {
        b->epoch.resize(n);
        b->roles.resize(n);
        for (int i = 0; i < n; i++) {
            b->epoch[i] = b->ctx[i]->epoch;
            b->roles[i] = roles_of(b->ctx[i]);
        }
        HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int rc = steps_in(fz, 2);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->st, &g);
        if (rc || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            for (int i = 0; i < n; i++) {
                rpmd_ctx* c = b->ctx[i];
                const Roles& R = b->roles[i];
                c->r = R.a[0], c->rst = R.a[1], c->dvdr = R.a[2], c->dvdrl = R.a[3], c->tmp = R.a[4];
            }
            return rc ? rc : RPMD_ERR_HIP;
        }
        if (hipGraphInstantiate(ge, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            return RPMD_ERR_HIP;
        }
        *gr = g;
        return 0;
    };
    // one untimed and one timed replay (two steps each) of a graph: real steps of the run
    auto timed = [&](hipGraphExec_t ge, float* ms) -> int {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return RPMD_ERR_HIP;
        int rc = 0;
        if (hipGraphLaunch(ge, b->st) != hipSuccess || hipEventRecord(t0, b->st) != hipSuccess ||
            hipGraphLaunch(ge, b->st) != hipSuccess || hipEventRecord(t1, b->st) != hipSuccess ||
            hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(ms, t0, t1) != hipSuccess)
            rc = RPMD_ERR_HIP;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        for (int& x : left) x -= 4;
        return rc;
    };
    if (nsteps >= 2) {
        bool valid = b->gexec != nullptr && b->fused == fused;
        for (int i = 0; i < n && valid; i++) valid = b->epoch[i] == b->ctx[i]->epoch;
        if (!valid) {
            batch_drop_graph(b);
            // one eager step of every replica first (lists, keys and launch records up to date
            // outside the capture), then two steps of every replica captured
            RETCHK(steps_in(fused, 1));
            for (int& x : left) x--;
            if (nsteps - 1 >= 2) {
                RETCHK(capture(fused, &b->graph, &b->gexec));
                b->fused = fused;
                // the batch autotuner, when a batched step is possible and enough steps are asked
                // for: four steps of each form, the second two of each timed, the faster kept
                if (eligible && !b->mode_tuned && nsteps - 1 >= 8) {
                    RETCHK(timed(b->gexec, &b->ms_fused));
                    hipGraph_t g2 = nullptr;
                    hipGraphExec_t ge2 = nullptr;
                    RETCHK(capture(false, &g2, &ge2));
                    const int rc = timed(ge2, &b->ms_separate);
                    b->mode_tuned = true;
                    b->mode_fused = b->ms_fused <= b->ms_separate;
                    if (!b->mode_fused) {
                        batch_drop_graph(b);
                        b->graph = g2;
                        b->gexec = ge2;
                        b->fused = false;
                    } else {
                        (void)hipGraphExecDestroy(ge2);
                        (void)hipGraphDestroy(g2);
                    }
                    RETCHK(rc);
                }
            }
        }
        if (b->gexec) {
            // replicas stepped an odd number of times since the capture (solo calls): one eager
            // step each brings their roles back
            std::vector<int> flip(n, 0);
            bool any = false;
            for (int i = 0; i < n; i++)
                if (roles_of(b->ctx[i]) != b->roles[i] && left[i] > 0) flip[i] = 1, any = true;
            if (any) {
                RETCHK(fork_steps(flip));
                for (int i = 0; i < n; i++) left[i] -= flip[i];
            }
            bool ok = true;
            for (int i = 0; i < n; i++) ok = ok && roles_of(b->ctx[i]) == b->roles[i];
            if (!ok) batch_drop_graph(b);
        }
        if (b->gexec) {
            int m = *std::min_element(left.begin(), left.end());
            for (; m >= 2; m -= 2) {
                HIPCHK(hipGraphLaunch(b->gexec, b->st));
                for (int& x : left) x -= 2;
            }
        }
    }
    RETCHK(fork_steps(left));
    HIPCHK(hipStreamSynchronize(b->st));
    for (rpmd_ctx* c : b->ctx) {
        HIPCHK(hipStreamSynchronize(c->st));
        RETCHK(run_status(c));
    }
    return 0;
}

int rpmd_batch_tuning(const rpmd_batch* b, int* batched, double* ms_batched, double* ms_separate, int* pair_cap,
                      int* pair_beside) {
    if (!b) return RPMD_ERR_ARG;
    if (batched) *batched = b->gexec ? (b->fused ? 1 : 0) : -1;
    if (ms_batched) *ms_batched = b->ms_fused;
    if (ms_separate) *ms_separate = b->ms_separate;
    if (pair_cap) *pair_cap = b->bpc;
    if (pair_beside) *pair_beside = b->two;
    return 0;
}

int rpmd_prof_begin(rpmd_ctx* c) {
    if (!c) return RPMD_ERR_ARG;
    c->prof = 1;
    c->ev_marks.clear();
    return 0;
}

int rpmd_prof_end(rpmd_ctx* c, double ms[RPMD_PROF_N], int counts[RPMD_PROF_N]) {
    if (!c) return RPMD_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->st));
    for (int k = 0; k < RPMD_PROF_N; k++) {
        ms[k] = 0.0;
        counts[k] = 0;
    }
    for (auto& m : c->ev_marks) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        ms[m.first] += t;
        counts[m.first] += 1;
    }
    c->prof = 0;
    c->ev_marks.clear();
    return 0;
}

int rpmd_estimators(rpmd_ctx* c, double out[RPMD_EST_N]) {
    if (!c || !c->st_valid || !out) return RPMD_ERR_ARG;
    const int na = c->na, nb = c->st_nb, nm = na / 3;
    const double beta = c->st_beta, wn = (double)nb / (beta * 1.0);
    // parts in tmp (scratch between steps), the ten sums in ev (50 doubles)
    const int nblk = estimator_blocks(na);
    launch_estimators(c->p, c->r, c->dvdr, c->dvdr2, c->mass, na, nb, wn * wn, c->tmp, c->st);
    ReduceJobs jobs;
    jobs.add(c->tmp, nblk, 0, 1.0);
    launch_reduce_ev(jobs, c->ev, c->st);
    double h[10];
    HIPCHK(hipMemcpyAsync(h, c->ev, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double en[4], vir[9];
    if (c->prm.rigid) {
        RETCHK(step_rig_energies(c, &en[0], vir, nullptr));
    } else {
        RETCHK(step_energies(c, &en[0], &en[1], &en[2], &en[3], vir, nullptr));
    }
    const double v = en[0];
    // /constraint/ nctot: three constraints per molecule when rigid (src/md_rattle.f90:29)
    const double nctot = c->prm.rigid ? 3.0 * nm : 0.0;
    const double tk = h[0], vrp = h[1];
    // virial_ke (src/kinetic.f90:42-49): nc = 3
    const double tq1x = 0.5 * h[2] / nb, tq1y = 0.5 * h[3] / nb, tq1z = 0.5 * h[4] / nb;
    const double tq2x = 0.5 * h[5] / nb, tq2y = 0.5 * h[6] / nb, tq2z = 0.5 * h[7] / nb;
    const double tq1 = 0.5 * (h[2] + h[3] + h[4]) / nb, tq2 = 0.5 * (h[5] + h[6] + h[7]) / nb;
    const double tc = 0.5 * (3.0 * na - 3.0 - nctot) / beta;
    out[RPMD_EST_KINETIC] = tk;
    out[RPMD_EST_SPRINGS] = vrp;
    out[RPMD_EST_POTENTIAL] = v;
    out[RPMD_EST_HRP] = ((double)nb * v + tk + vrp) / na;
    out[RPMD_EST_TEMP] = 3.157733e+5 * (2.0 * tk / ((3.0 * na - nctot - 3.0) * nb)) / nb;
    out[RPMD_EST_TV] = tc + tq1 + tq2;
    out[RPMD_EST_TQ1] = tq1;
    out[RPMD_EST_TQ2] = tq2;
    out[RPMD_EST_TVX] = tc / 3 + tq1x + tq2x;
    out[RPMD_EST_TVY] = tc / 3 + tq1y + tq2y;
    out[RPMD_EST_TVZ] = tc / 3 + tq1z + tq2z;
    return 0;
}

int rpmd_state_download(rpmd_ctx* c, double* p, double* r, double* dvdr, double* dvdr2, double energies[4],
                        double vir[9]) {
    if (!c || !c->st_valid) return RPMD_ERR_ARG;
    const int nb = c->st_nb;
    if (p) RETCHK(download_aos(c, c->p, p, nb));
    if (r) RETCHK(download_aos(c, c->r, r, nb));
    if (dvdr) RETCHK(download_aos(c, c->dvdr, dvdr, nb));
    if (dvdr2) RETCHK(download_aos(c, c->dvdr2, dvdr2, nb));
    if (energies && c->prm.rigid) {
        // v = vew + vlj of the forces(iopt = 1) call, vint = 0
        double h[60];
        HIPCHK(hipMemcpyAsync(h, c->ev_step, sizeof h, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        RETCHK(step_rig_energies(c, &energies[0], vir, nullptr));
        energies[2] = h[20];
        energies[1] = energies[0] - energies[2];
        energies[3] = 0.0;
    } else if (energies) {
        RETCHK(step_energies(c, &energies[0], &energies[1], &energies[2], &energies[3], vir, nullptr));
    }
    return 0;
}

}  // extern "C"
