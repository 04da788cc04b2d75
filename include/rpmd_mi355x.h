/*
 * rpmd_mi355x.h -- C-ABI of the MI355X (gfx950) ring-polymer force-and-propagate library.
 *
 * Every entry point takes plain pointers and sizes.  Arrays follow the reference's Fortran
 * layout: r(3,na,nb), dvdr(3,na,nb), p(3,na,nb) are column-major FP64 (xyz fastest, then
 * atom, then bead); virial(3,3) is column-major FP64.  Molecules are (O,H,H) triplets.
 * Atomic units (bohr, Hartree, electron masses, a.u. time) throughout.
 *
 * The caller owns every host array.  Inputs are never written (the reference overwrites the
 * oxygen with the M-site in place and restores it, src/md_forces.f90:181-188,213; this
 * library never writes r).  Device copies live in the context.  One context per host
 * thread; calls on a context are serialised and synchronous on return.
 *
 * Return value: RPMD_OK (0) or one of the RPMD_ERR_* codes.  rpmd_strerror() gives the
 * reference's own STOP message for the code.
 */
#ifndef RPMD_MI355X_H
#define RPMD_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: one per reference STOP condition on the path ---------------------- */
#define RPMD_OK                    0
#define RPMD_ERR_NA_NOT_MULT3      1  /* 'potenl 1'                 src/md_forces.f90:157-158 */
#define RPMD_ERR_NEIGH_OVERFLOW    2  /* 'neighbour list too small' src/neigh_list.f90:63    */
#define RPMD_ERR_NBDF_EVEN         3  /* 'Beadiabatic factor must be ODD' src/evolve_pi.f90:344-347 */
#define RPMD_ERR_NBDF_TOO_LARGE    4  /* 'must not exceed nb'       src/setup_options.f90:55-58 */
#define RPMD_ERR_RATTLE            5  /* 'max iterations exceeded'  src/md_rattle.f90:149-151   */
#define RPMD_ERR_REALFT_N          6  /* 'realft 1' (n > 256)       src/fourier.f90:40          */
#define RPMD_ERR_REALFT_MODE       7  /* 'realft 3'                 src/fourier.f90:56          */
#define RPMD_ERR_EWC_CUTOFF        8  /* 'ewc_basic : rcut too large' src/correct_tem.f90:104-109 */
#define RPMD_ERR_SMEARED           9  /* TEM correction (iopt 5) with smeared charges (wm,wh>0) */
#define RPMD_ERR_KMAX             10  /* Ewald kmax beyond the device k tiles (kmax > 31)      */
#define RPMD_ERR_ARG              11  /* bad argument (null pointer, nb/na beyond ctx sizes)  */
#define RPMD_ERR_HIP              12  /* HIP runtime error                                     */
#define RPMD_ERR_IOPT             13  /* iopt not in {0,1,2,3,4,5}                             */
#define RPMD_ERR_NOMEM            14  /* device allocation failed                              */
#define RPMD_ERR_THERM            15  /* thermostat not supported by this entry point          */
#define RPMD_ERR_NOISE            16  /* host noise buffer too short for the draws needed      */
#define RPMD_ERR_GLE              17  /* therm GLE without rpmd_gle_init for this (na, nb)     */
#define RPMD_ERR_TTAU             18  /* a thermostat is selected but ttau <= 0               */
#define RPMD_ERR_SOCKET           19  /* i-PI socket: no server at the address, or a broken stream */
#define RPMD_ERR_IPI_NAT          20  /* i-PI POSDATA with another atom count than the context's */
#define RPMD_ERR_BARO             21  /* barostat neither BER nor MCI                          */

/* ---- precision modes ------------------------------------------------------------------ */
#define RPMD_PREC_FP64   0  /* everything FP64 (the "plain FP64 GPU form")                   */
#define RPMD_PREC_MIXED  1  /* FP32 pair/erfc/eikr arithmetic, FP64 accumulators, FP64 state */

/* ---- intramolecular model switch ------------------------------------------------------ */
#define RPMD_INTRA_AUTO        0  /* reference rule src/md_forces.f90:254-263               */
#define RPMD_INTRA_HARMONIC    1  /* intra_harmonic (Toukan-Rahman SPC/F constants)         */
#define RPMD_INTRA_SPCFW       2  /* intra_spcfw (kb, ktheta harmonic bend)                 */
#define RPMD_INTRA_MORSE_HARM  3  /* intra_morse_harm (q-TIP4P/F)                           */
#define RPMD_INTRA_MORSE       4  /* intra_morse                                            */

/* ---- thermostats (3-letter codes of src/main.f90:218-252) ----------------------------- */
#define RPMD_THERM_NONE  0  /* NVE ('AND' and unknown codes do nothing inside evolve_pi)     */
#define RPMD_THERM_PRO   1  /* PILE-L   src/module_gle.f90:8                                 */
#define RPMD_THERM_PRM   2  /* PILE-G   src/module_gle.f90:75                                */
#define RPMD_THERM_PRG   3  /* Bussi global src/module_thermostat.f90:6                      */
#define RPMD_THERM_PRL   4  /* local Langevin src/module_thermostat.f90:56                   */
#define RPMD_THERM_GLE   5  /* GLE src/module_gle.f90:229                                    */

/*
 * Model parameters: the reference's COMMON blocks and thermostat module state, made explicit
 * (src/main.f90:53-72; src/module_thermostat.f90:1-4).
 */
typedef struct rpmd_params {
    /* /oo_param/ : LJ or Buckingham O-O term.  rcut already min(rcut, Lmin/2), bohr. */
    double oo_eps, oo_sig, oo_gam, rcut;
    /* /ew_param/ : alpha is the M-site fraction (1 = no M-site); wm, wh smeared widths. */
    double alpha, ecut, wm, wh;
    /* /ew_setup/ : filled by rpmd_setup_ewald (unit-box units when iamcub). */
    double wrcut, walpha, rkmax;
    int kmax;
    /* /symmetry/ */
    int iamcub;
    /* /intra_param/ and /geometry/ (theta in radians, reoh in bohr) */
    double apot, bpot, alp, alpb;
    double theta, reoh;
    /* /correct/ : TEM switch width, bohr (0 = no electrostatic contraction correction) */
    double sig;
    /* /beaddiabatic/ and /multiple_ts/ */
    int nbdf1, nbdf2;
    int mts;
    /* intramolecular model switch (RPMD_INTRA_*) */
    int intra_model;
    /* thermostat: RPMD_THERM_* and ttau in a.u. time */
    int therm;
    double ttau;
    /* /structure/ iamrigid: rigid molecules held by RATTLE (rpmd_evolve_rig_pi) */
    int rigid;
    /* RATTLE tolerance: relative on the squared bond lengths (stage 1), absolute on the bond
     * velocity multiplier (stage 2); <= 0 selects the reference's 1e-6 (src/md_rattle.f90:60,182) */
    double rat_tol;
} rpmd_params;

typedef struct rpmd_ctx rpmd_ctx;

/* Host-side Ewald set-up, src/setup_options.f90:104-189: fills wrcut, walpha, rkmax, kmax
 * and iamcub (iamcub on entry requests the cubic code, as the namelist flag does). */
int rpmd_setup_ewald(int na, const double boxlxyz[3], rpmd_params* prm);

/* Context: device buffers for up to na atoms x nb_max beads, a stream, cached tables.
 * precision is RPMD_PREC_FP64 or RPMD_PREC_MIXED.  No reference counterpart: the reference
 * keeps this state in COMMON blocks and reallocates scratch on every forces call
 * (src/md_forces.f90:19). */
int rpmd_ctx_create(const rpmd_params* prm, int na, int nb_max, int precision, rpmd_ctx** out);
int rpmd_ctx_set_params(rpmd_ctx* ctx, const rpmd_params* prm);
int rpmd_ctx_set_precision(rpmd_ctx* ctx, int precision);
int rpmd_ctx_destroy(rpmd_ctx* ctx);
const char* rpmd_strerror(int code);

/* subroutine forces(r,v,dvdr,nb,na,boxlxyz,z,virial,iopt)      src/md_forces.f90:1-110
 * iopt 0 all terms, 1 Ewald+LJ, 2 Ewald, 3 LJ, 4 intramolecular, 5 TEM Ewald correction.
 * v and virial are bead averages; dvdr is per bead (dV_k/dr_k). */
int rpmd_forces(rpmd_ctx* ctx, const double* r, int nb, int na, const double boxlxyz[3],
                const double* z, int iopt, double* v, double* dvdr, double* virial);

/* subroutine full_forces(r,na,nb,v,vew,voo,vint,vir,z,boxlxyz,dvdr,dvdr2)
 *                                                              src/md_forces.f90:305-472
 * dvdr = low-frequency (Ewald + O-O, RP-contracted when nbdf1/nbdf2 > 0), dvdr2 = intra. */
int rpmd_full_forces(rpmd_ctx* ctx, const double* r, int na, int nb, double* v, double* vew,
                     double* voo, double* vint, double* vir, const double* z,
                     const double boxlxyz[3], double* dvdr, double* dvdr2);

/* subroutine realft(data,m,n,mode)                              src/fourier.f90:16-63
 * In place: each of the m rows data(k,1:n) is R2HC (mode +1) or HC2R (mode -1)
 * transformed and scaled by 1/sqrt(n) (FFTW half-complex order). */
int rpmd_realft(rpmd_ctx* ctx, double* data, int m, int n, int mode);

/* Allocation-free variant of rpmd_realft.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, transform and errors as rpmd_realft, but no context: the transform matrix and the
 * staged rows live in a fixed device array of 8.9 MB that the library holds on each GPU it runs on
 * (rows are processed in chunks of 524288 / n, so m is unbounded; n > 256 returns
 * RPMD_ERR_REALFT_N).  It calls no hipMalloc or hipFree.  Every call on one GPU uses that GPU's
 * array: calls that overlap on one GPU (from several host threads) race. */
int rpmd_realft_noalloc(double* data, int m, int n, int mode);

/* The free ring-polymer update inside evolve_pi (src/evolve_pi.f90:109-125) with the monodromy of
 * ring_rpmd (src/md_pathi.f90:190-231) for time step dt: p and r(3,na,nb) are normal-mode
 * coordinates (realft mode +1 of the bead arrays) and are updated in place,
 * p' = p m1 + r m2, r' = p m3 + r m4 per mode, atom type mod(j,3) and component, every product
 * and sum rounded once in FP64: the same bits as the reference's loop compiled without FMA
 * contraction.  This is the update every evolve_pi step of this library applies. */
int rpmd_free_ring(rpmd_ctx* ctx, double* p, double* r, const double* mass, int na, int nb, double dt,
                   double beta);

/* subroutine evolve_pi(p,r,v,vew,vlj,vint,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,vir_lf,
 *                      irun,nbaro)                               src/evolve_pi.f90:1-322
 * One RPMD step (type 'RPMD', nbaro = 0).  p, r, dvdr, dvdr2 are updated in place.
 * noise: Gaussian deviates consumed in the reference's draw order (host RNG replay), or
 * NULL to draw them on the device from the context's counter-based stream. */
int rpmd_evolve_pi(rpmd_ctx* ctx, double* p, double* r, double* v, double* vew, double* vlj,
                   double* vint, double* dvdr, double* dvdr2, double dt, const double* mass,
                   int na, int nb, const double boxlxyz[3], const double* z, double beta,
                   double* vir, double* vir_lf, const double* noise, long long n_noise);

/* module barostat (src/module_barostat.f90:1-5; set up at src/main.f90:257-279): the barostat of
 * evolve_pi steps with nbaro = 1.  baro: RPMD_BARO_BER (Berendsen: taub = 41.341373 taufs, the
 * coupling time in a.u.) or RPMD_BARO_MCI (Monte Carlo volume moves: taub = taufs, one attempt per
 * taub steps on average; dv the largest volume change in bohr^3, the reference's 200); patm the
 * pressure in a.u. (pres / 2.9421e8 bar).  Resets the MC counters nacc, nmove. */
#define RPMD_BARO_NONE 0
#define RPMD_BARO_BER  1
#define RPMD_BARO_MCI  2
int rpmd_set_barostat(rpmd_ctx* ctx, int baro, double taub, double patm, double dv);
int rpmd_barostat_counts(const rpmd_ctx* ctx, long long* nacc, long long* nmove);

/* subroutine evolve_pi(...,irun,nbaro) with nbaro = 1              src/evolve_pi.f90:163-189
 * As rpmd_evolve_pi, and after the multiple time steps the barostat of rpmd_set_barostat acts on
 * the bead positions and the box (boxlxyz in/out), which the low-frequency forces of the step then
 * see.  vir_lf in: the previous step's low-frequency virial (vir = vir_lf + the step's mean
 * high-frequency virial drives BER); out: this step's.  BER: isotropic pressure in a cubic box
 * (src/module_barostat.f90:87-122), per axis otherwise (:28-85; the reference's beren_baro_nc reads
 * a local patm that hides the module's, :34, so its target is undefined there: the module's patm is
 * used here).  MCI (:124-220): uniforms are the ran2 draws in the reference's order (u in [0,1),
 * then, when u < 1/taub, v in [-1,1) and w in [0,1)), n_uniforms of them; NULL draws from the
 * library's own stream.  Every molecule moves by its center_water sum (the sum, not the mean, of
 * its three atoms' bead-averaged positions, src/md_tools.f90:367-390) times (scale - 1). */
int rpmd_evolve_pi_npt(rpmd_ctx* ctx, double* p, double* r, double* v, double* vew, double* vlj,
                       double* vint, double* dvdr, double* dvdr2, double dt, const double* mass,
                       int na, int nb, double boxlxyz[3], const double* z, double beta,
                       double* vir, double* vir_lf, const double* noise, long long n_noise,
                       const double* uniforms, int n_uniforms);

/* subroutine evolve_rig_pi(p,r,v,v_lf,v_hf,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,vir_lf,
 *                          irun,nbaro)                          src/evolve_rigid.f90:100-234
 * One rigid-molecule RPMD step (params.rigid = 1, nbaro = 0): nb bead-spring sub-steps with
 * RATTLE (src/md_rattle.f90, FP64, tolerance params.rat_tol, <= 0 the reference's 1e-6, at most
 * 1000 iterations), forces(iopt = 1), and RATTLE stage 2 with dt.  vir is in/out: the reference
 * adds the incoming value.  Thermostats PRG and PRL as the reference's; PRO (PILE-L on the normal
 * modes) and GLE, which the reference's rigid step does not offer, as the same half-step kicks
 * followed by a projection of the momenta (and the GLE auxiliary momenta) onto the constraint
 * surface; other codes leave the step thermostat-free, as there.  A RATTLE that does not converge
 * returns RPMD_ERR_RATTLE. */
int rpmd_evolve_rig_pi(rpmd_ctx* ctx, double* p, double* r, double* v, double* v_lf, double* v_hf,
                       double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                       const double boxlxyz[3], const double* z, double beta, double* vir,
                       double* vir_lf, const double* noise, long long n_noise);

/* subroutine evolve_rig_pi(...,irun,nbaro) with nbaro = 1          src/evolve_rigid.f90:203-220
 * As rpmd_evolve_rig_pi, then the barostat of rpmd_set_barostat acts on the final positions and the
 * box (boxlxyz in/out).  BER: the pressure of the step's in/out virial vir and of virial_ke with the
 * step's intermolecular gradient (nctot = 3 nm); the molecules move rigidly.  MCI: as mc_baro, which
 * replaces v (v_lf), vir, dvdr and dvdr2 by its full_forces values when it tries a move.  uniforms
 * as in rpmd_evolve_pi_npt. */
int rpmd_evolve_rig_pi_npt(rpmd_ctx* ctx, double* p, double* r, double* v, double* v_lf, double* v_hf,
                           double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                           double boxlxyz[3], const double* z, double beta, double* vir,
                           double* vir_lf, const double* noise, long long n_noise,
                           const double* uniforms, int n_uniforms);

/* subroutine sample_rpmd(p,na,nb,mass,beta,irun)                 src/module_thermostat.f90:199-226
 * Maxwell-Boltzmann bead momenta at beta/nb, then pshift (:229-264) removes the total momentum.
 * noise: 3*na*nb standard normal deviates in the reference's draw order (bead, atom, xyz), or NULL
 * to draw them on the device from seed.  p is written in the reference layout p(3,na,nb). */
int rpmd_sample_rpmd(rpmd_ctx* ctx, double* p, int na, int nb, const double* mass, double beta,
                     const double* noise, long long n_noise, unsigned long long seed);

/* subroutine therm_gle_init(ttau,na,nb,dt,irun,beta)              src/module_gle.f90:146-227
 * Sets up the GLE thermostat (therm = RPMD_THERM_GLE) of this context for (na, nb): A is the
 * (ns+1) x (ns+1) drift matrix in the GLE-A file's row order (A[i*(ns+1)+j] = A(i+1,j+1)), in
 * units of 2 pi / ttau (ttau from the context parameters); C = (nb/beta) I.  dt is the time step
 * the reference passes here, half the evolve_pi step (src/main.f90:609).  gp: the initial
 * auxiliary momenta in the reference layout gp(3,na,nb,ns+1), or NULL to draw chol(C) xi on the
 * device from seed.  ns <= 7. */
int rpmd_gle_init(rpmd_ctx* ctx, int ns, const double* A, int na, int nb, double dt, double beta,
                  const double* gp, unsigned long long seed);

/* Device-resident stepping (no counterpart in the reference, whose trajectory loop calls
 * evolve once per step, src/md_dynamics.f90:499-517): upload a state once, run nsteps
 * evolve_pi steps (evolve_rig_pi steps when params.rigid) with all data in HBM, read
 * energies/state back when needed.  energies = {v, vew, vlj, vint}. */
int rpmd_state_upload(rpmd_ctx* ctx, const double* p, const double* r, const double* mass,
                      const double* z, int na, int nb, const double boxlxyz[3], double beta,
                      double dt, unsigned long long seed);
int rpmd_run_steps(rpmd_ctx* ctx, int nsteps);
int rpmd_state_download(rpmd_ctx* ctx, double* p, double* r, double* dvdr, double* dvdr2,
                        double energies[4], double vir[9]);
/* Estimators of the device-resident state, computed on the device after rpmd_run_steps (no state
 * download): out[RPMD_EST_N].
 *   KINETIC   tk = sum over beads and atoms of |p|^2 / (2 m)       kinetic_rpmd, src/kinetic.f90:248-265
 *   SPRINGS   vrp = sum m wn^2 |r_k - r_{k-1}|^2 / 2, wn = nb/beta  src/md_eq.f90:257-274
 *   POTENTIAL v, the bead-averaged potential of the last step
 *   HRP       (nb v + tk + vrp) / na, the conserved quantity of NVE RPMD per atom  src/md_eq.f90:278
 *   TEMP      toK 2 tk / ((3 na - nctot - 3) nb) / nb, kelvin      src/md_eq.f90:276
 *   TV, TQ1, TQ2, TVX..TVZ  the centroid-virial kinetic energy of virial_ke with the state's low-
 *             and high-frequency gradients (src/kinetic.f90:1-54; nc = 3, nctot = 3 nm when rigid) */
#define RPMD_EST_KINETIC   0
#define RPMD_EST_SPRINGS   1
#define RPMD_EST_POTENTIAL 2
#define RPMD_EST_HRP       3
#define RPMD_EST_TEMP      4
#define RPMD_EST_TV        5
#define RPMD_EST_TQ1       6
#define RPMD_EST_TQ2       7
#define RPMD_EST_TVX       8
#define RPMD_EST_TVY       9
#define RPMD_EST_TVZ      10
#define RPMD_EST_N        11
int rpmd_estimators(rpmd_ctx* ctx, double out[RPMD_EST_N]);

/* Capture rpmd_run_steps(ctx, 1) into a HIP graph and replay it for later calls. */
int rpmd_use_graph(rpmd_ctx* ctx, int enable);
/* Streams of a device-resident force call (no counterpart in the reference): 2 (default) runs the
 * real-space Ewald branch on a second stream beside the k-space kernels; 3 also runs the O-O list
 * and pair sums, with the intramolecular forces of a step, on a third; 1 runs every kernel on one
 * stream, as when kernels are timed.  Results are the same either way. */
int rpmd_ctx_set_streams(rpmd_ctx* ctx, int streams);

/* Replica batches (no counterpart in the reference, which runs trajectories one after another):
 * contexts with uploaded states (each with its own na, nb, box, Ewald set-up, temperature,
 * thermostat; no RP contraction) advanced together, nsteps steps each, by one hipGraph.  When
 * every replica takes the fused PILE-L step in mixed precision (cubic box, mts 1, up to 64 beads),
 * the step can be batched: the real-space pair walk, k_sk_h16, k_sk_final and k_kforce_h16 of all
 * replicas as one launch each, the rest on the replicas' streams; otherwise, and when a batch
 * autotuner found it faster, each replica takes its own step side by side.  The autotuner runs
 * once, in the first call of nine steps or more: four graph-replayed steps of each form (real
 * steps: both forms give the same bits), the faster kept.  Re-create the batch after a context's
 * state is uploaded again. */
typedef struct rpmd_batch rpmd_batch;
int rpmd_batch_create(rpmd_ctx* const* ctxs, int n, rpmd_batch** out);
int rpmd_batch_run_steps(rpmd_batch* batch, int nsteps);
int rpmd_batch_destroy(rpmd_batch* batch);
/* The autotuners' choices (no reference counterpart): batched (1) or separate (0) steps in use,
 * -1 before the first step; the timed pair of steps of each form (ms, 0 if not measured); for the
 * batched form the pair walk's cap (blocks per CU, 0 none) and whether it runs beside the k-space
 * kernels (1) or after them (0). */
int rpmd_batch_tuning(const rpmd_batch* batch, int* batched, double* ms_batched, double* ms_separate, int* pair_cap,
                      int* pair_beside);

/* subroutine run_driver(p,r,dvdr,dvdr2,na,nb,boxlxyz,z,beta,dt,mass,irun,itst,pt,pb,print)
 *                                                        src/ipi_driver.f90:1-254, src/sockets.c
 * The i-PI client: connects to the server at address ("prefix:host:port", as the reference reads
 * its serverdata file: prefix "unix" gives the UNIX socket /tmp/ipi_<host>, any other an internet
 * socket to host:port) and serves it until it disconnects or sends anything but STATUS, POSDATA
 * or GETFORCE: STATUS -> READY or HAVEDATA; POSDATA (cell, its inverse, the atom count, bohr
 * positions) -> full_forces on one bead of this context with the box the cell's diagonal;
 * GETFORCE -> FORCEREADY, the potential, the atom count, the forces -(dvdr + dvdr2), minus the
 * transposed virial, an empty extra field.  z: the charges.  served (optional): GETFORCE replies.
 * A POSDATA atom count other than na returns RPMD_ERR_IPI_NAT (the reference prints a warning
 * and reads past its arrays). */
int rpmd_ipi_run(rpmd_ctx* ctx, const char* address, const double* z, int na, int* served);

/* Kernel timing (no reference counterpart; the reference has no timers): HIP events recorded on
 * the context stream around the Ewald kernels of eager launches between begin and end.
 * ms/counts index: 0 real-space pair kernel, 1 structure-factor GEMM, 2 k-space force kernel,
 * 3 structure-factor weights/energy kernel, 4 real-space list upkeep (the rebuild test and, when
 * due, the rebuild). */
#define RPMD_PROF_N 5
int rpmd_prof_begin(rpmd_ctx* ctx);
int rpmd_prof_end(rpmd_ctx* ctx, double ms[RPMD_PROF_N], int counts[RPMD_PROF_N]);

#ifdef __cplusplus
}
#endif
#endif /* RPMD_MI355X_H */
