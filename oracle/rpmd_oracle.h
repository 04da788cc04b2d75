/*
 * rpmd_oracle.h -- CPU FP64 restatement of the reference RPMD force-and-propagate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the timed CPU
 * baseline -- never as part of the product path.
 *
 * Arrays use the reference's Fortran layout (column-major r(3,na,nb), vir(3,3)).
 * The parameter struct is the product's rpmd_params (include/rpmd_mi355x.h), which is
 * a plain restatement of the reference COMMON blocks.
 */
#ifndef RPMD_ORACLE_H
#define RPMD_ORACLE_H
#include "../include/rpmd_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Numerical-Recipes ran2 state + the cached Box-Muller deviate (src/random.f90:1-72). */
typedef struct orc_rng {
    int idum, idum2, iy, iv[32];
    int iset;
    double gset;
} orc_rng;

typedef struct orc_gle {
    int ns, na, nb;
    double gT[64], gS[64];   /* (ns+1)^2 column-major, ns <= 7 */
    double* gp;              /* gp(3,na,nb,ns+1) */
} orc_gle;

int  orc_setup_ewald(int na, const double box[3], rpmd_params* prm);
void orc_setup_box_size(int nm, double rho, double wmass, const int ncell[3], double box[3]);

int orc_forces(const rpmd_params* P, const double* r, int nb, int na, const double box[3],
               const double* z, int iopt, double* v, double* dvdr, double* vir);
int orc_full_forces(const rpmd_params* P, const double* r, int na, int nb, double* v,
                    double* vew, double* voo, double* vint, double* vir, const double* z,
                    const double box[3], double* dvdr, double* dvdr2);
int orc_rp_contract_nm(const rpmd_params* P, const double* rnm, double* v, double* dvdr,
                       int nb, int na, const double box[3], const double* z, double* vir,
                       int nbdf, int iopt);

int  orc_realft(double* data, int m, int n, int mode);
void orc_ring_rpmd(const double* mass, double* monod, int na, int nb, double dt, double beta);
void orc_free_ring(double* p, double* r, const double* monod, int na, int nb);

void   orc_rng_init(orc_rng* s);
void   orc_rng_seed(orc_rng* s, int irun);
double orc_ran2(orc_rng* s, double rmin, double rmax);
double orc_gaussian(orc_rng* s, double sigma);
void   orc_gaussian_fill(orc_rng* s, double* out, long long n);

void orc_therm_pro(double* p, const double* mass, double beta, double ctau, double dt,
                   int na, int nb, orc_rng* s);
void orc_therm_pro_global(double* p, const double* mass, double beta, double ctau, double dt,
                          int na, int nb, orc_rng* s);
void orc_parinello_therm(double* p, const double* mass, double tau, int na, int nb,
                         double dt, orc_rng* s, double beta, int nctot);
void orc_parinello_therm_loc(double* p, const double* mass, double tau, int na, int nb,
                             double dt, orc_rng* s, double beta);
void orc_sample_rpmd(double* p, int na, int nb, const double* mass, double beta, orc_rng* s);
void orc_pshift(double* p, const double* mass, int na, int nb);

int  orc_gle_init(orc_gle* g, const double* A, int ns, double ttau, int na, int nb,
                  double dt, orc_rng* s, double beta);
void orc_gle_free(orc_gle* g);
void orc_therm_gle(orc_gle* g, double* p, const double* mass, int na, int nb, orc_rng* s);
void orc_matrix_exp(const double* M, int n, int j, int k, double* EM);
void orc_cholesky(const double* SST, double* S, int n);

/* module barostat (src/module_barostat.f90:1-5): baro, taub, patm, dv and the MC counters */
#define ORC_BARO_BER 1
#define ORC_BARO_MCI 2
typedef struct orc_baro {
    int baro;
    double taub, patm, dv;
    int nacc, nmove;
} orc_baro;
void   orc_virial_ke(const double* r, const double* dvdr, const double* dvdr2, double* tv, double tvxyz[3],
                     double* tq1, double* tq2, double beta, int na, int nb, int nctot);
double orc_pres_lj_tail(const rpmd_params* P, const double box[3], int na);
double orc_pressure(const rpmd_params* P, const double vir[9], double tv, int na, const double box[3]);
void   orc_center_water(const double* r, double* pc, int nm, int nb);
void   orc_beren_driver(const rpmd_params* P, const orc_baro* B, const double vir[9], double tv,
                        const double tvxyz[3], double dt, double* r, double box[3], int na, int nb);
int    orc_mc_baro(const rpmd_params* P, orc_baro* B, double* r, double* dvdr, double* dvdr2, double* vir,
                   double* eo, const double* z, double beta, double box[3], int na, int nb, orc_rng* s);
int orc_evolve_pi_npt(const rpmd_params* P, double* p, double* r, double* v, double* vew, double* vlj,
                      double* vint, double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                      double box[3], const double* z, double beta, double* vir, double* vir_lf, orc_rng* s,
                      orc_gle* g, orc_baro* B);

int orc_evolve_pi(const rpmd_params* P, double* p, double* r, double* v, double* vew,
                  double* vlj, double* vint, double* dvdr, double* dvdr2, double dt,
                  const double* mass, int na, int nb, const double box[3], const double* z,
                  double beta, double* vir, double* vir_lf, orc_rng* s, orc_gle* g);

int orc_evolve_rig_pi_npt(const rpmd_params* P, double* p, double* r, double* v, double* v_lf, double* v_hf,
                          double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb, double box[3],
                          const double* z, double beta, double* vir, double* vir_lf, orc_rng* s, orc_baro* B);

int orc_rattle_s1(double* r, const double* rold, double* p, int nm, const double* mass, double dt,
                  double* dvdr, double* vir, double reoh, double theta);
int orc_rattle_s2(const double* r, double* p, int nm, const double* mass, double dt, double* dvdr,
                  double* vir, double reoh, double theta);
void orc_bead_forces(const double* r, int na, int nb, double* dvdrb, double beta, const double* mass);
int orc_evolve_rig_pi(const rpmd_params* P, double* p, double* r, double* v, double* v_lf, double* v_hf,
                      double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                      const double box[3], const double* z, double beta, double* vir, double* vir_lf,
                      orc_rng* s);

#ifdef __cplusplus
}
#endif
#endif
