/*
 * rpmd_oracle.c -- CPU FP64 restatement of the reference RPMD force-and-propagate path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); never linked into the
 * product.  Built with gcc -O2 -ffp-contract=off, no fast-math.
 *
 * Each function names the reference routine (file:line) whose semantics it restates.  Loop
 * orders follow the reference where they fix the summation order.  Index convention: the
 * reference's 1-based atom numbers are kept inside the linked-cell and neighbour-list code
 * (0 terminates a chain, as in the reference); arrays are addressed 0-based.
 *
 * Parity status of each part (how it is pinned) is listed in tests/ and DESIGN.md:
 *   forces (all iopt)       -- vs the reference's own Fortran built from source (oracle/_ref)
 *                              and the gentest bead frames (tests/golden)
 *   realft                  -- closed-form DFT identities (FFTW itself is absent)
 *   evolve_pi (NVE)         -- gentest 20-step NVE windows (tests/golden)
 *   ran2/gaussian, PILE, GLE helpers -- vs oracle/_ref (random.f90, module_gle.f90)
 */
#include "rpmd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXNAB 200 /* src/globals.inc:294 */

static double nint_d(double x) { return round(x); } /* Fortran NINT: half away from zero */
#define RX(r, c, j) (r)[3 * (j) + (c)]               /* r(c+1, j+1) of a one-bead array */
#define VIR(v, i, j) (v)[3 * (j) + (i)]              /* vir(i+1, j+1), column-major     */

/* ======================================================================================
 * Linked cells: src/linked_cell.f90
 * ==================================================================================== */
typedef struct {
    int nc[3], mc[3], nn;
    int* link; /* nc0*nc1*nc2, 1-based head atom, 0 = empty */
    int* list; /* list[i] for 1-based atom i, 0 = end       */
    int ink[62][3];
    int* map[3]; /* map[d][(ic-1)*(2m+1) + (jc+m)] = 1-based cell */
} cells_t;

static void cells_free(cells_t* C) {
    free(C->link);
    free(C->list);
    for (int d = 0; d < 3; d++) free(C->map[d]);
}

/* cell_setup src/linked_cell.f90:1-32 */
static void cell_setup(const double box[3], double rcut, int mc[3], int nc[3], int* nn) {
    int n = 1;
    for (int i = 0; i < 3; i++) {
        if (3.0 * rcut > box[i]) {
            nc[i] = 1;
            mc[i] = 0;
        } else {
            int c = (int)(2.0 * box[i] / rcut);
            nc[i] = c > 6 ? c : 6;
            mc[i] = 2;
        }
        n *= 2 * mc[i] + 1;
    }
    *nn = (n - 1) / 2;
}

/* cell_map_nc src/linked_cell.f90:188-282 (cell_map :131 is the cubic special case) */
static void cell_map_nc(cells_t* C) {
    int k = 0;
    for (int iz = 1; iz <= C->mc[2]; iz++)
        for (int iy = -C->mc[1]; iy <= C->mc[1]; iy++)
            for (int ix = -C->mc[0]; ix <= C->mc[0]; ix++) {
                C->ink[k][0] = ix; C->ink[k][1] = iy; C->ink[k][2] = iz; k++;
            }
    for (int iy = 1; iy <= C->mc[1]; iy++)
        for (int ix = -C->mc[0]; ix <= C->mc[0]; ix++) {
            C->ink[k][0] = ix; C->ink[k][1] = iy; C->ink[k][2] = 0; k++;
        }
    for (int ix = 1; ix <= C->mc[0]; ix++) {
        C->ink[k][0] = ix; C->ink[k][1] = 0; C->ink[k][2] = 0; k++;
    }
    C->nn = k;
    for (int d = 0; d < 3; d++) {
        int m = C->mc[d], n = C->nc[d], w = 2 * m + 1;
        C->map[d] = (int*)malloc(sizeof(int) * n * w);
        for (int ic = 1; ic <= n; ic++)
            for (int jc = -m; jc <= m; jc++) {
                int kc = ic + jc;
                if (kc < 1) kc += n;
                else if (kc > n) kc -= n;
                C->map[d][(ic - 1) * w + (jc + m)] = kc;
            }
    }
}

static int cmap(const cells_t* C, int d, int ic, int jc) {
    return C->map[d][(ic - 1) * (2 * C->mc[d] + 1) + (jc + C->mc[d])];
}
static int* clink(const cells_t* C, int ix, int iy, int iz) {
    return &C->link[((iz - 1) * C->nc[1] + (iy - 1)) * C->nc[0] + (ix - 1)];
}

/* cell_list src/linked_cell.f90:34-89 (atoms 1, 1+ijump, ...) */
static void cell_list(const double* r, int n, const double box[3], int ijump, cells_t* C) {
    C->link = (int*)calloc((size_t)C->nc[0] * C->nc[1] * C->nc[2], sizeof(int));
    C->list = (int*)calloc((size_t)n + 1, sizeof(int));
    double cell[3], half[3], onbox[3];
    for (int d = 0; d < 3; d++) {
        onbox[d] = 1.0 / box[d];
        half[d] = 0.5 * box[d];
        cell[d] = (double)C->nc[d] / box[d];
    }
    for (int i = 1; i <= n; i += ijump) {
        int ic[3];
        for (int d = 0; d < 3; d++) {
            double dx = RX(r, d, i - 1) - box[d] * nint_d(onbox[d] * RX(r, d, i - 1));
            int c = 1 + (int)((dx + half[d]) * cell[d]);
            ic[d] = c < C->nc[d] ? c : C->nc[d];
        }
        int* h = clink(C, ic[0], ic[1], ic[2]);
        C->list[i] = *h;
        *h = i;
    }
    cell_map_nc(C);
}

/* cell_list_unit src/linked_cell.f90:91-129: unit box, ncell per side, mcell = 0 or 2 */
static void cell_list_unit(const double* r, int n, int ncell, int mcell, cells_t* C) {
    for (int d = 0; d < 3; d++) {
        C->nc[d] = ncell;
        C->mc[d] = mcell;
    }
    C->link = (int*)calloc((size_t)ncell * ncell * ncell, sizeof(int));
    C->list = (int*)calloc((size_t)n + 1, sizeof(int));
    double cell = (double)ncell;
    for (int i = 1; i <= n; i++) {
        int ic[3];
        for (int d = 0; d < 3; d++) {
            double dx = RX(r, d, i - 1) - nint_d(RX(r, d, i - 1));
            long long c = 1 + (long long)((dx + 0.5) * cell);
            ic[d] = (int)(c < ncell ? c : ncell);
        }
        int* h = clink(C, ic[0], ic[1], ic[2]);
        C->list[i] = *h;
        *h = i;
    }
    cell_map_nc(C);
    if (mcell == 0) C->nn = 0;
}

typedef void (*pair_fn)(int i, int j, void* ctx); /* 1-based atoms */

/* rwald / rwald_cell order: every cell's own pairs first, then the half-shell neighbours
 * (src/ewald_cubic.f90:114-230, src/ewald_noncubic.f90:108-224). */
static void cells_pairs_global(const cells_t* C, pair_fn fn, void* ctx) {
    for (int iz = 1; iz <= C->nc[2]; iz++)
        for (int iy = 1; iy <= C->nc[1]; iy++)
            for (int ix = 1; ix <= C->nc[0]; ix++) {
                int i = *clink(C, ix, iy, iz);
                while (i > 0) {
                    int j = C->list[i];
                    while (j > 0) { fn(i, j, ctx); j = C->list[j]; }
                    i = C->list[i];
                }
            }
    for (int k = 0; k < C->nn; k++) {
        int incx = C->ink[k][0], incy = C->ink[k][1], incz = C->ink[k][2];
        for (int iz = 1; iz <= C->nc[2]; iz++) {
            int jz = cmap(C, 2, iz, incz);
            for (int iy = 1; iy <= C->nc[1]; iy++) {
                int jy = cmap(C, 1, iy, incy);
                for (int ix = 1; ix <= C->nc[0]; ix++) {
                    int jx = cmap(C, 0, ix, incx);
                    int i = *clink(C, ix, iy, iz);
                    while (i > 0) {
                        int j = *clink(C, jx, jy, jz);
                        while (j > 0) { fn(i, j, ctx); j = C->list[j]; }
                        i = C->list[i];
                    }
                }
            }
        }
    }
}

/* lj_cell / buck_cell / ewc_cell_tem order: per cell, own pairs then its half-shell
 * neighbours (src/ljones.f90:258-362). */
static void cells_pairs_percell(const cells_t* C, pair_fn fn, void* ctx) {
    for (int iz = 1; iz <= C->nc[2]; iz++)
        for (int iy = 1; iy <= C->nc[1]; iy++)
            for (int ix = 1; ix <= C->nc[0]; ix++) {
                int i = *clink(C, ix, iy, iz);
                while (i > 0) {
                    int j = C->list[i];
                    while (j > 0) { fn(i, j, ctx); j = C->list[j]; }
                    i = C->list[i];
                }
                for (int k = 0; k < C->nn; k++) {
                    int jx = cmap(C, 0, ix, C->ink[k][0]);
                    int jy = cmap(C, 1, iy, C->ink[k][1]);
                    int jz = cmap(C, 2, iz, C->ink[k][2]);
                    int i2 = *clink(C, ix, iy, iz);
                    while (i2 > 0) {
                        int j = *clink(C, jx, jy, jz);
                        while (j > 0) { fn(i2, j, ctx); j = C->list[j]; }
                        i2 = C->list[i2];
                    }
                }
            }
}

/* ======================================================================================
 * Neighbour lists: src/neigh_list.f90 (CSR: point(na+3), list(maxnab*na), 1-based)
 * ==================================================================================== */
typedef struct {
    int* point; /* 1-based index: point[j] for atom j, size na+4 */
    int* list;  /* 1-based: list[1..nlist] */
    int cap;
} nlist_t;

static int nlist_alloc(nlist_t* L, int na) {
    L->point = (int*)calloc((size_t)na + 4, sizeof(int));
    L->cap = MAXNAB * na;
    L->list = (int*)calloc((size_t)L->cap + 2, sizeof(int));
    return 0;
}
static void nlist_free(nlist_t* L) { free(L->point); free(L->list); }

static void mimage(const double box[3], const double onbox[3], double* d) {
    for (int c = 0; c < 3; c++) d[c] = d[c] - box[c] * nint_d(onbox[c] * d[c]);
}

/* nei_list src/neigh_list.f90:23-75 */
static int nei_list(const double* r, int na, nlist_t* L, int njump, double rcut, const double box[3]) {
    double rcutsq = rcut * rcut, onbox[3];
    for (int c = 0; c < 3; c++) onbox[c] = 1.0 / box[c];
    memset(L->point, 0, sizeof(int) * (na + 4));
    memset(L->list, 0, sizeof(int) * (L->cap + 2));
    int nl = 0, jend = na - njump + 1;
    for (int j = 1; j <= jend; j += njump) {
        L->point[j] = nl + 1;
        for (int i = j + njump; i <= na; i += njump) {
            double d[3];
            for (int c = 0; c < 3; c++) d[c] = RX(r, c, i - 1) - RX(r, c, j - 1);
            mimage(box, onbox, d);
            double drsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (drsq < rcutsq) {
                nl++;
                if (nl == L->cap) return RPMD_ERR_NEIGH_OVERFLOW;
                L->list[nl] = i;
            }
        }
    }
    for (int i = jend + njump; i <= na + njump; i += njump) {
        nl++;
        L->point[i] = nl;
    }
    return 0;
}

/* nei_cell src/neigh_list.f90:77-220 */
static int nei_cell(const double* r, int na, nlist_t* L, int njump, double rcut, const double box[3]) {
    cells_t C;
    memset(&C, 0, sizeof C);
    cell_setup(box, rcut, C.mc, C.nc, &C.nn);
    cell_list(r, na, box, njump, &C);
    double onbox[3], half[3], cell[3], rcutsq = rcut * rcut;
    for (int c = 0; c < 3; c++) {
        onbox[c] = 1.0 / box[c];
        half[c] = 0.5 * box[c];
        cell[c] = C.nc[c] / box[c];
    }
    memset(L->point, 0, sizeof(int) * (na + 4));
    memset(L->list, 0, sizeof(int) * (L->cap + 2));
    int nl = 0, jend = na - njump + 1, err = 0;
    for (int j = 1; j <= jend && !err; j += njump) {
        L->point[j] = nl + 1;
        int jc[3];
        for (int c = 0; c < 3; c++) {
            double dx = RX(r, c, j - 1) - box[c] * nint_d(onbox[c] * RX(r, c, j - 1));
            int q = 1 + (int)((dx + half[c]) * cell[c]);
            jc[c] = q < C.nc[c] ? q : C.nc[c];
        }
        int i = *clink(&C, jc[0], jc[1], jc[2]);
        while (i != j) i = C.list[i];
        i = C.list[i];
        while (i > 0) {
            double d[3];
            for (int c = 0; c < 3; c++) d[c] = RX(r, c, i - 1) - RX(r, c, j - 1);
            mimage(box, onbox, d);
            double drsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (drsq < rcutsq) {
                nl++;
                if (nl == L->cap) { err = RPMD_ERR_NEIGH_OVERFLOW; break; }
                L->list[nl] = i;
            }
            i = C.list[i];
        }
        for (int k = 0; k < C.nn && !err; k++) {
            int ix = cmap(&C, 0, jc[0], C.ink[k][0]);
            int iy = cmap(&C, 1, jc[1], C.ink[k][1]);
            int iz = cmap(&C, 2, jc[2], C.ink[k][2]);
            i = *clink(&C, ix, iy, iz);
            while (i > 0) {
                double d[3];
                for (int c = 0; c < 3; c++) d[c] = RX(r, c, i - 1) - RX(r, c, j - 1);
                mimage(box, onbox, d);
                double drsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (drsq < rcutsq) {
                    nl++;
                    if (nl == L->cap) { err = RPMD_ERR_NEIGH_OVERFLOW; break; }
                    L->list[nl] = i;
                }
                i = C.list[i];
            }
        }
    }
    if (!err)
        for (int i = jend + njump; i <= na + njump; i += njump) {
            nl++;
            L->point[i] = nl;
        }
    cells_free(&C);
    return err;
}

/* nei_driver src/neigh_list.f90:1-21 */
static int nei_driver(const double* r, int na, nlist_t* L, int njump, double rcut, const double box[3]) {
    double bmax = box[0];
    if (box[1] > bmax) bmax = box[1];
    if (box[2] > bmax) bmax = box[2];
    if (3.0 * rcut > bmax) return nei_list(r, na, L, njump, rcut, box);
    return nei_cell(r, na, L, njump, rcut, box);
}

/* ======================================================================================
 * Ewald set-up: src/setup_options.f90:104-189, box: :191-270
 * ==================================================================================== */
int orc_setup_ewald(int na, const double box[3], rpmd_params* P) {
    double pi = acos(-1.0);
    if (box[0] == box[1] && box[0] == box[2] && P->iamcub) {
        P->iamcub = 1;
        double cvar = 1.2;
        double w = cvar * pow((double)na, -1.0 / 6.0);
        P->wrcut = w < 0.5 ? w : 0.5;
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
        double cvar = 1.2;
        double rshort = box[0], rlong = box[0];
        for (int d = 1; d < 3; d++) {
            if (box[d] < rshort) rshort = box[d];
            if (box[d] > rlong) rlong = box[d];
        }
        double w = cvar * pow((double)na, -1.0 / 6.0);
        P->wrcut = rshort * (w < 0.5 ? w : 0.5);
        P->walpha = pi / P->wrcut;
        P->rkmax = 2.0 * pi * P->walpha;
        P->kmax = (int)(P->walpha * rlong);
    }
    return 0;
}

void orc_setup_box_size(int nm, double rho, double wmass, const int ncell[3], double box[3]) {
    const double toA = 0.5291772108, emass = 9.1093826e-31;
    double vbox = (nm * emass * wmass) / (rho * 1000.0);
    vbox = vbox / pow(toA * 1e-10, 3);
    double totcell = (double)ncell[0] * ncell[1] * ncell[2];
    double boxl = pow(vbox / totcell, 1.0 / 3.0);
    for (int d = 0; d < 3; d++) box[d] = boxl * ncell[d];
}

/* ======================================================================================
 * Real-space Ewald with the A&S 7.1.26 erfc (src/ewald_cubic.f90:78-233;
 * src/ewald_noncubic.f90:54-312)
 * ==================================================================================== */
#define AS_P 3.0525860
#define AS_A1 0.254829592
#define AS_A2 -0.284496736
#define AS_A3 1.421413741
#define AS_A4 -1.453152027
#define AS_A5 1.061405429

typedef struct {
    const double *r, *q;
    const int* mol;
    double* dvdr;
    double v, vir[9];
    double rcutsq, alpha, sfac;
    int unit;         /* 1: unit box, nint(dx) image; 0: physical box */
    double box[3], onbox[3];
    int smeared;      /* Gaussian charges: wid = {M-M, H-H, M-H} inverse widths */
    double wid[3];
} rw_ctx;

static void rwald_pair(int i1, int j1, void* vctx) {
    rw_ctx* c = (rw_ctx*)vctx;
    int i = i1 - 1, j = j1 - 1;
    double dx = RX(c->r, 0, i) - RX(c->r, 0, j);
    double dy = RX(c->r, 1, i) - RX(c->r, 1, j);
    double dz = RX(c->r, 2, i) - RX(c->r, 2, j);
    if (c->unit) {
        dx = dx - nint_d(dx);
        dy = dy - nint_d(dy);
        dz = dz - nint_d(dz);
    } else {
        dx = dx - c->box[0] * nint_d(dx * c->onbox[0]);
        dy = dy - c->box[1] * nint_d(dy * c->onbox[1]);
        dz = dz - c->box[2] * nint_d(dz * c->onbox[2]);
    }
    double drsq = dx * dx + dy * dy + dz * dz;
    if (drsq < c->rcutsq) {
        double dr = sqrt(drsq);
        double x = c->alpha * dr;
        double e = exp(-x * x);
        double t = AS_P / (AS_P + x);
        double erfc_ = e * t * (AS_A1 + t * (AS_A2 + t * (AS_A3 + t * (AS_A4 + t * AS_A5))));
        double sfacs = 0.0, es = 0.0;
        if (c->mol[i] == c->mol[j]) {
            erfc_ = erfc_ - 1.0;
        } else if (c->smeared) {
            /* rwald_smeared src/ewald_smeared.f90:71-94: the first site of a molecule is type 1 (M),
             * the others type 2 (H); like types take wid(type), unlike ones wid(3) */
            int ti = (i % 3) ? 2 : 1, tj = (j % 3) ? 2 : 1;
            double wt = ti == tj ? c->wid[ti - 1] : c->wid[2];
            double xs = dr * wt;
            sfacs = wt * (-2.0 / sqrt(acos(-1.0)));
            es = exp(-xs * xs);
            double ts = AS_P / (AS_P + xs);
            double erfcs = es * ts * (AS_A1 + ts * (AS_A2 + ts * (AS_A3 + ts * (AS_A4 + ts * AS_A5))));
            erfc_ = erfc_ - erfcs;
        }
        double du = erfc_ / dr;
        double dur = (c->sfac * e - sfacs * es - du) / drsq;
        double qij = c->q[i] * c->q[j];
        double dv = qij * du, dvr = qij * dur;
        double dvx = dvr * dx, dvy = dvr * dy, dvz = dvr * dz;
        c->v = c->v + dv;
        RX(c->dvdr, 0, i) += dvx; RX(c->dvdr, 1, i) += dvy; RX(c->dvdr, 2, i) += dvz;
        RX(c->dvdr, 0, j) -= dvx; RX(c->dvdr, 1, j) -= dvy; RX(c->dvdr, 2, j) -= dvz;
        VIR(c->vir, 0, 0) += dx * dvx; VIR(c->vir, 0, 1) += dx * dvy; VIR(c->vir, 0, 2) += dx * dvz;
        VIR(c->vir, 1, 0) += dy * dvx; VIR(c->vir, 1, 1) += dy * dvy; VIR(c->vir, 1, 2) += dy * dvz;
        VIR(c->vir, 2, 0) += dz * dvx; VIR(c->vir, 2, 1) += dz * dvy; VIR(c->vir, 2, 2) += dz * dvz;
    }
}

/* kwald src/ewald_cubic.f90:235-533 (unit box; loop structure kept for the k ordering) */
static void kwald(int n, const double* q, const double* r, double* v, double* vir, double* dvdr,
                  double alpha, double rkmax2, int kmax) {
    size_t K = (size_t)kmax + 1;
    double* cx = (double*)malloc(sizeof(double) * 2 * n * K);
    double* cy = (double*)malloc(sizeof(double) * 2 * n * K);
    double* cz = (double*)malloc(sizeof(double) * 2 * n * K);
    double* cj = (double*)malloc(sizeof(double) * 4 * n);
    double* ck = (double*)malloc(sizeof(double) * 8 * n);
#define CX(a, i, k) cx[((size_t)(k) * n + (i)) * 2 + (a)]
#define CY(a, i, k) cy[((size_t)(k) * n + (i)) * 2 + (a)]
#define CZ(a, i, k) cz[((size_t)(k) * n + (i)) * 2 + (a)]
    double pi = acos(-1.0), rlat = 2.0 * pi;
    for (int i = 0; i < n; i++) {
        CX(0, i, 0) = q[i]; CX(1, i, 0) = 0.0;
        double arg = rlat * RX(r, 0, i), c = cos(arg), s = sin(arg);
        for (int k = 1; k <= kmax; k++) {
            CX(0, i, k) = c * CX(0, i, k - 1) - s * CX(1, i, k - 1);
            CX(1, i, k) = s * CX(0, i, k - 1) + c * CX(1, i, k - 1);
        }
    }
    for (int i = 0; i < n; i++) {
        CY(0, i, 0) = 1.0; CY(1, i, 0) = 0.0;
        double arg = rlat * RX(r, 1, i), c = cos(arg), s = sin(arg);
        for (int k = 1; k <= kmax; k++) {
            CY(0, i, k) = c * CY(0, i, k - 1) - s * CY(1, i, k - 1);
            CY(1, i, k) = s * CY(0, i, k - 1) + c * CY(1, i, k - 1);
        }
    }
    for (int i = 0; i < n; i++) {
        CZ(0, i, 0) = 1.0; CZ(1, i, 0) = 0.0;
        double arg = rlat * RX(r, 2, i), c = cos(arg), s = sin(arg);
        for (int k = 1; k <= kmax; k++) {
            CZ(0, i, k) = c * CZ(0, i, k - 1) - s * CZ(1, i, k - 1);
            CZ(1, i, k) = s * CZ(0, i, k - 1) + c * CZ(1, i, k - 1);
        }
    }
    double b = 0.25 / (alpha * alpha), f = rlat;
    double V = *v;
    for (int kx = 0; kx <= kmax; kx++) {
        if (kx == 1) f = 2.0 * f;
        double rkx = rlat * kx, rky, rkz, rk2 = rkx * rkx;
        if (rk2 > 0.0) {
            double s1 = 0, s2 = 0;
            for (int i = 0; i < n; i++) { s1 = s1 + CX(0, i, kx); s2 = s2 + CX(1, i, kx); }
            double w = (f / rk2) * exp(-b * rk2);
            double et = w * (s1 * s1 + s2 * s2);
            V = V + et;
            double term = 2.0 * (1.0 / rk2 + b);
            VIR(vir, 0, 0) += et * (term * rkx * rkx - 1.0);
            VIR(vir, 1, 1) += et * (-1.0);
            VIR(vir, 2, 2) += et * (-1.0);
            w = w + w;
            double wkx = w * rkx;
            for (int i = 0; i < n; i++) {
                double t2 = s2 * CX(0, i, kx) - s1 * CX(1, i, kx);
                RX(dvdr, 0, i) += wkx * t2;
            }
        }
        for (int kz = 1; kz <= kmax; kz++) {
            rky = 0.0;
            rkz = rlat * kz;
            rk2 = rkx * rkx + rkz * rkz;
            if (rk2 > rkmax2) break;
            double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            for (int i = 0; i < n; i++) {
                double cc = CX(0, i, kx) * CZ(0, i, kz), cs = CX(0, i, kx) * CZ(1, i, kz);
                double sc = CX(1, i, kx) * CZ(0, i, kz), ss = CX(1, i, kx) * CZ(1, i, kz);
                cj[4 * i + 0] = cc - ss; cj[4 * i + 1] = sc + cs;
                cj[4 * i + 2] = cc + ss; cj[4 * i + 3] = sc - cs;
                s1 = s1 + cj[4 * i + 0]; s2 = s2 + cj[4 * i + 1];
                s3 = s3 + cj[4 * i + 2]; s4 = s4 + cj[4 * i + 3];
            }
            double w = (f / rk2) * exp(-b * rk2);
            double et = w * (s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4);
            V = V + et;
            double term = 2.0 * (1.0 / rk2 + b);
            double xx = et * (term * rkx * rkx - 1.0), xz = et * (term * rkx * rkz);
            double yy = et * (-1.0), zz = et * (term * rkz * rkz - 1.0);
            VIR(vir, 0, 0) += xx; VIR(vir, 0, 2) += xz; VIR(vir, 1, 1) += yy;
            VIR(vir, 2, 0) += xz; VIR(vir, 2, 2) += zz;
            w = w + w;
            double wkx = w * rkx, wkz = w * rkz;
            for (int i = 0; i < n; i++) {
                double t2 = s2 * cj[4 * i + 0] - s1 * cj[4 * i + 1];
                double t4 = s4 * cj[4 * i + 2] - s3 * cj[4 * i + 3];
                RX(dvdr, 0, i) += wkx * (t2 + t4);
                RX(dvdr, 2, i) += wkz * (t2 - t4);
            }
        }
        for (int ky = 1; ky <= kmax; ky++) {
            rkz = 0.0;
            rky = rlat * ky;
            rk2 = rkx * rkx + rky * rky;
            if (rk2 > rkmax2) break;
            double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            for (int i = 0; i < n; i++) {
                double cc = CX(0, i, kx) * CY(0, i, ky), cs = CX(0, i, kx) * CY(1, i, ky);
                double sc = CX(1, i, kx) * CY(0, i, ky), ss = CX(1, i, kx) * CY(1, i, ky);
                cj[4 * i + 0] = cc - ss; cj[4 * i + 1] = sc + cs;
                cj[4 * i + 2] = cc + ss; cj[4 * i + 3] = sc - cs;
                s1 = s1 + cj[4 * i + 0]; s2 = s2 + cj[4 * i + 1];
                s3 = s3 + cj[4 * i + 2]; s4 = s4 + cj[4 * i + 3];
            }
            double w = (f / rk2) * exp(-b * rk2);
            double et = w * (s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4);
            V = V + et;
            double term = 2.0 * (1.0 / rk2 + b);
            double xx = et * (term * rkx * rkx - 1.0), xy = et * (term * rkx * rky);
            double yy = et * (term * rky * rky - 1.0), zz = et * (-1.0);
            VIR(vir, 0, 0) += xx; VIR(vir, 0, 1) += xy; VIR(vir, 1, 0) += xy;
            VIR(vir, 1, 1) += yy; VIR(vir, 2, 2) += zz;
            w = w + w;
            double wkx = w * rkx, wky = w * rky;
            for (int i = 0; i < n; i++) {
                double t2 = s2 * cj[4 * i + 0] - s1 * cj[4 * i + 1];
                double t4 = s4 * cj[4 * i + 2] - s3 * cj[4 * i + 3];
                RX(dvdr, 0, i) += wkx * (t2 + t4);
                RX(dvdr, 1, i) += wky * (t2 - t4);
            }
            for (int kz = 1; kz <= kmax; kz++) {
                rkz = rlat * kz;
                rk2 = rkx * rkx + rky * rky + rkz * rkz;
                if (rk2 > rkmax2) break;
                double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (int i = 0; i < n; i++) {
                    double* a = &cj[4 * i];
                    double* o = &ck[8 * i];
                    double cca = a[0] * CZ(0, i, kz), csa = a[0] * CZ(1, i, kz);
                    double sca = a[1] * CZ(0, i, kz), ssa = a[1] * CZ(1, i, kz);
                    double ccb = a[2] * CZ(0, i, kz), csb = a[2] * CZ(1, i, kz);
                    double scb = a[3] * CZ(0, i, kz), ssb = a[3] * CZ(1, i, kz);
                    o[0] = cca - ssa; o[1] = sca + csa; o[2] = cca + ssa; o[3] = sca - csa;
                    o[4] = ccb - ssb; o[5] = scb + csb; o[6] = ccb + ssb; o[7] = scb - csb;
                    for (int m = 0; m < 8; m++) s[m] = s[m] + o[m];
                }
                double w = (f / rk2) * exp(-b * rk2);
                double et = w * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + s[3] * s[3] +
                                 s[4] * s[4] + s[5] * s[5] + s[6] * s[6] + s[7] * s[7]);
                V = V + et;
                double term = 2.0 * (1.0 / rk2 + b);
                double xx = et * (term * rkx * rkx - 1.0), xy = et * (term * rkx * rky);
                double xz = et * (term * rkx * rkz), yy = et * (term * rky * rky - 1.0);
                double yz = et * (term * rky * rkz), zz = et * (term * rkz * rkz - 1.0);
                VIR(vir, 0, 0) += xx; VIR(vir, 0, 1) += xy; VIR(vir, 0, 2) += xz;
                VIR(vir, 1, 0) += xy; VIR(vir, 1, 1) += yy; VIR(vir, 1, 2) += yz;
                VIR(vir, 2, 0) += xz; VIR(vir, 2, 1) += yz; VIR(vir, 2, 2) += zz;
                w = w + w;
                double wkx = w * rkx, wky = w * rky, wkz = w * rkz;
                for (int i = 0; i < n; i++) {
                    double* o = &ck[8 * i];
                    double t2 = s[1] * o[0] - s[0] * o[1];
                    double t4 = s[3] * o[2] - s[2] * o[3];
                    double t6 = s[5] * o[4] - s[4] * o[5];
                    double t8 = s[7] * o[6] - s[6] * o[7];
                    RX(dvdr, 0, i) += wkx * (t2 + t4 + t6 + t8);
                    RX(dvdr, 1, i) += wky * (t2 + t4 - t6 - t8);
                    RX(dvdr, 2, i) += wkz * (t2 - t4 + t6 - t8);
                }
            }
        }
    }
    double vs = 0.0;
    for (int i = 0; i < n; i++) vs = vs + q[i] * q[i];
    double rfac = alpha / sqrt(pi);
    vs = rfac * vs;
    *v = V - vs;
    free(cx); free(cy); free(cz); free(cj); free(ck);
#undef CX
#undef CY
#undef CZ
}

/* ewald_cubic src/ewald_cubic.f90:1-76 */
static void ewald_cubic(const rpmd_params* P, int n, const int* mol, const double* q,
                        const double* r, double* v, double* vir, double* dvdr, double boxl) {
    double* rs = (double*)malloc(sizeof(double) * 3 * n);
    double onboxl = 1.0 / boxl;
    for (int i = 0; i < 3 * n; i++) rs[i] = onboxl * r[i];
    double rkmax2 = P->rkmax * P->rkmax;
    int ncell, mcell;
    if (3.0 * P->wrcut > 1.0) {
        ncell = 1;
        mcell = 0;
    } else {
        int c = (int)(2.0 / P->wrcut);
        ncell = c > 6 ? c : 6;
        mcell = 2;
    }
    rw_ctx c;
    memset(&c, 0, sizeof c);
    c.r = rs; c.q = q; c.mol = mol; c.dvdr = dvdr;
    c.rcutsq = P->wrcut * P->wrcut;
    c.alpha = P->walpha;
    double pi = acos(-1.0);
    c.sfac = -2.0 * (P->walpha / sqrt(pi));
    c.unit = 1;
    if (P->wm > 0.0 || P->wh > 0.0) {
        /* smeared charges, widths in unit-box lengths (src/ewald_cubic.f90:40-54) */
        double sm = sqrt(1.0 / (2.0 * P->wm)) * onboxl, sh = sqrt(1.0 / (2.0 * P->wh)) * onboxl;
        c.smeared = 1;
        c.wid[0] = 1.0 / (2.0 * sm);
        c.wid[1] = 1.0 / (2.0 * sh);
        c.wid[2] = 1.0 / sqrt(2.0 * (sm * sm + sh * sh));
    }
    cells_t C;
    memset(&C, 0, sizeof C);
    cell_list_unit(rs, n, ncell, mcell, &C);
    cells_pairs_global(&C, rwald_pair, &c);
    cells_free(&C);
    double V = c.v;
    kwald(n, q, rs, &V, c.vir, dvdr, P->walpha, rkmax2, P->kmax);
    *v = onboxl * V;
    for (int i = 0; i < 9; i++) vir[i] = onboxl * c.vir[i];
    double onbox2 = onboxl * onboxl;
    for (int i = 0; i < 3 * n; i++) dvdr[i] = onbox2 * dvdr[i];
    free(rs);
}

/* kwald_nc src/ewald_noncubic.f90:314-457 */
static void kwald_nc(const double* r, const double* z, int n, double* v, double* vir,
                     double* dvdr, double alpha, double rkmax2, int kmax, const double box[3]) {
    int K = kmax, W = 2 * kmax + 1;
    double* ckx = (double*)malloc(sizeof(double) * n * (K + 1));
    double* skx = (double*)malloc(sizeof(double) * n * (K + 1));
    double* cky = (double*)malloc(sizeof(double) * n * W);
    double* sky = (double*)malloc(sizeof(double) * n * W);
    double* ckz = (double*)malloc(sizeof(double) * n * W);
    double* skz = (double*)malloc(sizeof(double) * n * W);
    double* cxy = (double*)malloc(sizeof(double) * n);
    double* sxy = (double*)malloc(sizeof(double) * n);
    double* dsdr = (double*)malloc(sizeof(double) * 6 * n);
#define IX(i, k) ((size_t)(k) * n + (i))
#define IY(i, k) ((size_t)((k) + K) * n + (i))
    double pi = acos(-1.0), rfac = alpha / sqrt(pi), twopi = 2.0 * pi;
    double xlat = twopi / box[0], ylat = twopi / box[1], zlat = twopi / box[2];
    for (int i = 0; i < n; i++) {
        ckx[IX(i, 0)] = z[i]; skx[IX(i, 0)] = 0.0;
        double xi = xlat * RX(r, 0, i), c = cos(xi), s = sin(xi);
        for (int k = 1; k <= K; k++) {
            ckx[IX(i, k)] = c * ckx[IX(i, k - 1)] - s * skx[IX(i, k - 1)];
            skx[IX(i, k)] = s * ckx[IX(i, k - 1)] + c * skx[IX(i, k - 1)];
        }
    }
    for (int i = 0; i < n; i++) {
        cky[IY(i, 0)] = 1.0; sky[IY(i, 0)] = 0.0;
        double yi = ylat * RX(r, 1, i), c = cos(yi), s = sin(yi);
        for (int k = 1; k <= K; k++) {
            cky[IY(i, k)] = c * cky[IY(i, k - 1)] - s * sky[IY(i, k - 1)];
            sky[IY(i, k)] = s * cky[IY(i, k - 1)] + c * sky[IY(i, k - 1)];
            cky[IY(i, -k)] = cky[IY(i, k)];
            sky[IY(i, -k)] = -sky[IY(i, k)];
        }
    }
    for (int i = 0; i < n; i++) {
        ckz[IY(i, 0)] = 1.0; skz[IY(i, 0)] = 0.0;
        double zi = zlat * RX(r, 2, i), c = cos(zi), s = sin(zi);
        for (int k = 1; k <= K; k++) {
            ckz[IY(i, k)] = c * ckz[IY(i, k - 1)] - s * skz[IY(i, k - 1)];
            skz[IY(i, k)] = s * ckz[IY(i, k - 1)] + c * skz[IY(i, k - 1)];
            ckz[IY(i, -k)] = ckz[IY(i, k)];
            skz[IY(i, -k)] = -skz[IY(i, k)];
        }
    }
    double b = 0.25 / (alpha * alpha), f = twopi / (box[0] * box[1] * box[2]);
    double V = *v;
    for (int kx = 0; kx <= K; kx++) {
        if (kx == 1) f = 2.0 * f;
        double rkx = xlat * kx;
        for (int ky = -K; ky <= K; ky++) {
            double rky = ylat * ky;
            for (int i = 0; i < n; i++) {
                cxy[i] = ckx[IX(i, kx)] * cky[IY(i, ky)] - skx[IX(i, kx)] * sky[IY(i, ky)];
                sxy[i] = ckx[IX(i, kx)] * sky[IY(i, ky)] + skx[IX(i, kx)] * cky[IY(i, ky)];
            }
            for (int kz = -K; kz <= K; kz++) {
                double rkz = zlat * kz;
                double rk2 = rkx * rkx + rky * rky + rkz * rkz;
                if (rk2 < rkmax2 && rk2 != 0.0) {
                    double sr = 0.0, si = 0.0;
                    for (int i = 0; i < n; i++) {
                        double tr = cxy[i] * ckz[IY(i, kz)] - sxy[i] * skz[IY(i, kz)];
                        double ti = cxy[i] * skz[IY(i, kz)] + sxy[i] * ckz[IY(i, kz)];
                        sr = sr + tr;
                        si = si + ti;
                        dsdr[6 * i + 0] = -rkx * ti; dsdr[6 * i + 1] = rkx * tr;
                        dsdr[6 * i + 2] = -rky * ti; dsdr[6 * i + 3] = rky * tr;
                        dsdr[6 * i + 4] = -rkz * ti; dsdr[6 * i + 5] = rkz * tr;
                    }
                    double w = (f / rk2) * exp(-b * rk2);
                    double et = w * (sr * sr + si * si);
                    V = V + et;
                    w = 2.0 * w;
                    sr = w * sr;
                    si = w * si;
                    for (int i = 0; i < n; i++) {
                        RX(dvdr, 0, i) += sr * dsdr[6 * i + 0] + si * dsdr[6 * i + 1];
                        RX(dvdr, 1, i) += sr * dsdr[6 * i + 2] + si * dsdr[6 * i + 3];
                        RX(dvdr, 2, i) += sr * dsdr[6 * i + 4] + si * dsdr[6 * i + 5];
                    }
                    double term = 2.0 * (1.0 / rk2 + b);
                    double xx = et * (term * rkx * rkx - 1.0), xy = et * (term * rkx * rky);
                    double xz = et * (term * rkx * rkz), yy = et * (term * rky * rky - 1.0);
                    double yz = et * (term * rky * rkz), zz = et * (term * rkz * rkz - 1.0);
                    VIR(vir, 0, 0) += xx; VIR(vir, 0, 1) += xy; VIR(vir, 0, 2) += xz;
                    VIR(vir, 1, 0) += xy; VIR(vir, 1, 1) += yy; VIR(vir, 1, 2) += yz;
                    VIR(vir, 2, 0) += xz; VIR(vir, 2, 1) += yz; VIR(vir, 2, 2) += zz;
                }
            }
        }
    }
    double vs = 0.0;
    for (int i = 0; i < n; i++) vs = vs + z[i] * z[i];
    vs = rfac * vs;
    *v = V - vs;
    free(ckx); free(skx); free(cky); free(sky); free(ckz); free(skz);
    free(cxy); free(sxy); free(dsdr);
#undef IX
#undef IY
}

/* ewald_nc src/ewald_noncubic.f90:1-52 */
static void ewald_nc(const rpmd_params* P, int n, const int* mol, const double* q,
                     const double* r, double* v, double* vir, double* dvdr, const double box[3]) {
    rw_ctx c;
    memset(&c, 0, sizeof c);
    c.r = r; c.q = q; c.mol = mol; c.dvdr = dvdr;
    c.rcutsq = P->wrcut * P->wrcut;
    c.alpha = P->walpha;
    double pi = acos(-1.0);
    c.sfac = -2.0 * (P->walpha / sqrt(pi));
    c.unit = 0;
    for (int d = 0; d < 3; d++) {
        c.box[d] = box[d];
        c.onbox[d] = 1.0 / box[d];
    }
    double bmax = box[0];
    if (box[1] > bmax) bmax = box[1];
    if (box[2] > bmax) bmax = box[2];
    if (P->wm > 0.0 || P->wh > 0.0) {
        /* smeared charges: rwald_smeared_nc, always over cells (src/ewald_noncubic.f90:18-33,
         * src/ewald_smeared.f90:219-453) */
        double sm = sqrt(1.0 / (2.0 * P->wm)), sh = sqrt(1.0 / (2.0 * P->wh));
        c.smeared = 1;
        c.wid[0] = 1.0 / (2.0 * sm);
        c.wid[1] = 1.0 / (2.0 * sh);
        c.wid[2] = 1.0 / sqrt(2.0 * (sm * sm + sh * sh));
        cells_t C;
        memset(&C, 0, sizeof C);
        cell_setup(box, P->wrcut, C.mc, C.nc, &C.nn);
        cell_list(r, n, box, 1, &C);
        cells_pairs_global(&C, rwald_pair, &c);
        cells_free(&C);
    } else if (3.0 * P->wrcut > bmax) {
        /* rwald_basic :229-312 */
        for (int i = 1; i <= n; i++)
            for (int j = 1; j <= i - 1; j++) rwald_pair(i, j, &c);
    } else {
        /* rwald_cell :54-227 */
        cells_t C;
        memset(&C, 0, sizeof C);
        cell_setup(box, P->wrcut, C.mc, C.nc, &C.nn);
        cell_list(r, n, box, 1, &C);
        cells_pairs_global(&C, rwald_pair, &c);
        cells_free(&C);
    }
    double V = c.v;
    kwald_nc(r, q, n, &V, c.vir, dvdr, P->walpha, P->rkmax * P->rkmax, P->kmax, box);
    *v = V;
    memcpy(vir, c.vir, sizeof(double) * 9);
}

/* ======================================================================================
 * O-O Lennard-Jones (src/ljones.f90) and Buckingham exp-6 (src/buckingham.f90)
 * ==================================================================================== */
typedef struct {
    const double* r;
    double* dvdr;
    double v, vir[9];
    double rcutsq, sigsq;
    double a, b, c; /* Buckingham */
    double box[3], onbox[3];
    int buck;
} oo_ctx;

/* pair body for lj_basic/lj_cell (dx = r_i - r_j, dvdr_i -= f dx) */
static void lj_pair(int i1, int j1, void* vctx) {
    oo_ctx* c = (oo_ctx*)vctx;
    int i = i1 - 1, j = j1 - 1;
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = RX(c->r, k, i) - RX(c->r, k, j);
    mimage(c->box, c->onbox, d);
    double dx = d[0], dy = d[1], dz = d[2];
    double drsq = dx * dx + dy * dy + dz * dz;
    if (drsq < c->rcutsq) {
        if (!c->buck) {
            double onr2 = 1.0 / drsq, sr2 = c->sigsq * onr2, sr6 = sr2 * sr2 * sr2;
            double vij = sr6 * (sr6 - 1.0);
            c->v = c->v + vij;
            double wij = sr6 * (sr6 - 0.5), fij = wij * onr2;
            double dfx = fij * dx, dfy = fij * dy, dfz = fij * dz;
            RX(c->dvdr, 0, i) -= dfx; RX(c->dvdr, 1, i) -= dfy; RX(c->dvdr, 2, i) -= dfz;
            RX(c->dvdr, 0, j) += dfx; RX(c->dvdr, 1, j) += dfy; RX(c->dvdr, 2, j) += dfz;
            VIR(c->vir, 0, 0) -= dx * dfx; VIR(c->vir, 1, 1) -= dy * dfy; VIR(c->vir, 2, 2) -= dz * dfz;
            VIR(c->vir, 0, 1) -= dx * dfy; VIR(c->vir, 0, 2) -= dx * dfz; VIR(c->vir, 1, 0) -= dy * dfx;
            VIR(c->vir, 1, 2) -= dy * dfz; VIR(c->vir, 2, 0) -= dz * dfx; VIR(c->vir, 2, 1) -= dz * dfy;
        } else {
            double br6 = c->b / (drsq * drsq * drsq), dr = sqrt(drsq);
            double exp6 = c->a * exp(c->c * dr);
            c->v = c->v + (exp6 - br6);
            double onr = 1.0 / dr, fij = c->c * exp6 + 6.0 * br6 * onr;
            double dfx = fij * dx * onr, dfy = fij * dy * onr, dfz = fij * dz * onr;
            RX(c->dvdr, 0, i) += dfx; RX(c->dvdr, 1, i) += dfy; RX(c->dvdr, 2, i) += dfz;
            RX(c->dvdr, 0, j) -= dfx; RX(c->dvdr, 1, j) -= dfy; RX(c->dvdr, 2, j) -= dfz;
            VIR(c->vir, 0, 0) += dx * dfx; VIR(c->vir, 1, 1) += dy * dfy; VIR(c->vir, 2, 2) += dz * dfz;
            VIR(c->vir, 0, 1) += dx * dfy; VIR(c->vir, 0, 2) += dx * dfz; VIR(c->vir, 1, 0) += dy * dfx;
            VIR(c->vir, 1, 2) += dy * dfz; VIR(c->vir, 2, 0) += dz * dfx; VIR(c->vir, 2, 1) += dz * dfy;
        }
    }
}

/* lj_driver src/ljones.f90:1-29 and buck_driver src/buckingham.f90:1-29 */
static void oo_driver(const rpmd_params* P, const double* r, double* dvdr, double* v, double* vir,
                      const nlist_t* L, int na, const double box[3], int njump) {
    oo_ctx c;
    memset(&c, 0, sizeof c);
    c.r = r; c.dvdr = dvdr;
    c.rcutsq = P->rcut * P->rcut;
    c.sigsq = P->oo_sig * P->oo_sig;
    c.buck = (P->oo_gam != 0.0);
    if (c.buck) {
        c.a = exp(P->oo_gam) * 6.0 * P->oo_eps / (P->oo_gam - 6.0);
        c.b = P->oo_eps * pow(P->oo_sig, 6) / (1.0 - (6.0 / P->oo_gam));
        c.c = -P->oo_gam / P->oo_sig;
    }
    for (int d = 0; d < 3; d++) { c.box[d] = box[d]; c.onbox[d] = 1.0 / box[d]; }
    memset(dvdr, 0, sizeof(double) * 3 * na);
    if (L && L->list[1] != 0) {
        /* lj_list :111-209 / buck_list :103-195: dx = r_i - r_j with j the list owner */
        for (int j = 1; j <= na; j += njump) {
            int ibeg = L->point[j], iend = L->point[j + njump] - 1;
            if (ibeg <= iend) {
                double dj[3] = {RX(dvdr, 0, j - 1), RX(dvdr, 1, j - 1), RX(dvdr, 2, j - 1)};
                for (int inab = ibeg; inab <= iend; inab++) {
                    int i = L->list[inab];
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = RX(r, k, i - 1) - RX(r, k, j - 1);
                    mimage(box, c.onbox, d);
                    double dx = d[0], dy = d[1], dz = d[2];
                    double drsq = dx * dx + dy * dy + dz * dz;
                    if (drsq < c.rcutsq) {
                        if (!c.buck) {
                            double onr2 = 1.0 / drsq, sr2 = c.sigsq * onr2, sr6 = sr2 * sr2 * sr2;
                            c.v = c.v + sr6 * (sr6 - 1.0);
                            double fij = sr6 * (sr6 - 0.5) * onr2;
                            double dfx = fij * dx, dfy = fij * dy, dfz = fij * dz;
                            RX(dvdr, 0, i - 1) -= dfx; RX(dvdr, 1, i - 1) -= dfy; RX(dvdr, 2, i - 1) -= dfz;
                            dj[0] += dfx; dj[1] += dfy; dj[2] += dfz;
                            VIR(c.vir, 0, 0) -= dx * dfx; VIR(c.vir, 1, 1) -= dy * dfy; VIR(c.vir, 2, 2) -= dz * dfz;
                            VIR(c.vir, 0, 1) -= dx * dfy; VIR(c.vir, 0, 2) -= dx * dfz; VIR(c.vir, 1, 0) -= dy * dfx;
                            VIR(c.vir, 1, 2) -= dy * dfz; VIR(c.vir, 2, 0) -= dz * dfx; VIR(c.vir, 2, 1) -= dz * dfy;
                        } else {
                            double br6 = c.b / (drsq * drsq * drsq), dr = sqrt(drsq);
                            double exp6 = c.a * exp(c.c * dr);
                            c.v = c.v + (exp6 - br6);
                            double onr = 1.0 / dr, fij = c.c * exp6 + 6.0 * br6 * onr;
                            double dfx = fij * dx * onr, dfy = fij * dy * onr, dfz = fij * dz * onr;
                            RX(dvdr, 0, i - 1) += dfx; RX(dvdr, 1, i - 1) += dfy; RX(dvdr, 2, i - 1) += dfz;
                            dj[0] -= dfx; dj[1] -= dfy; dj[2] -= dfz;
                            VIR(c.vir, 0, 0) += dx * dfx; VIR(c.vir, 1, 1) += dy * dfy; VIR(c.vir, 2, 2) += dz * dfz;
                            VIR(c.vir, 0, 1) += dx * dfy; VIR(c.vir, 0, 2) += dx * dfz; VIR(c.vir, 1, 0) += dy * dfx;
                            VIR(c.vir, 1, 2) += dy * dfz; VIR(c.vir, 2, 0) += dz * dfx; VIR(c.vir, 2, 1) += dz * dfy;
                        }
                    }
                }
                RX(dvdr, 0, j - 1) = dj[0]; RX(dvdr, 1, j - 1) = dj[1]; RX(dvdr, 2, j - 1) = dj[2];
            }
        }
    } else {
        double bmax = box[0];
        if (box[1] > bmax) bmax = box[1];
        if (box[2] > bmax) bmax = box[2];
        if (3.0 * P->rcut > bmax) {
            /* lj_basic :32-109: j outer from 1+njump, i inner below it */
            for (int j = 1 + njump; j <= na; j += njump)
                for (int i = 1; i <= j - njump; i += njump) lj_pair(i, j, &c);
        } else {
            /* lj_cell :211-379 */
            cells_t C;
            memset(&C, 0, sizeof C);
            cell_setup(box, P->rcut, C.mc, C.nc, &C.nn);
            cell_list(r, na, box, njump, &C);
            cells_pairs_percell(&C, lj_pair, &c);
            cells_free(&C);
        }
    }
    if (!c.buck) {
        double vscale = 4.0 * P->oo_eps, dscale = 48.0 * P->oo_eps;
        c.v = vscale * c.v;
        for (int k = 0; k < 9; k++) c.vir[k] = dscale * c.vir[k];
        for (int i = 1; i <= na; i += njump)
            for (int k = 0; k < 3; k++) RX(dvdr, k, i - 1) = dscale * RX(dvdr, k, i - 1);
    }
    *v = c.v;
    memcpy(vir, c.vir, sizeof(double) * 9);
}

/* ======================================================================================
 * Intramolecular terms (src/intra_harmonic.f90, intra_spcfw.f90, intra_morse_harm.f90,
 * intra_morse.f90).  All accumulate into dvdr; v and vir are set.
 * ==================================================================================== */
typedef struct {
    double dx[3][3]; /* bond vectors d1=H1-O, d2=H2-O, d3=H2-H1 */
    double dr[3];    /* lengths */
    double u[3][3];  /* unit vectors */
} geom_t;

static void mol_geom(const double* r, int j, geom_t* g) {
    for (int c = 0; c < 3; c++) {
        g->dx[0][c] = RX(r, c, j + 1) - RX(r, c, j);
        g->dx[1][c] = RX(r, c, j + 2) - RX(r, c, j);
        g->dx[2][c] = RX(r, c, j + 2) - RX(r, c, j + 1);
    }
    for (int b = 0; b < 3; b++) {
        double sq = g->dx[b][0] * g->dx[b][0] + g->dx[b][1] * g->dx[b][1] + g->dx[b][2] * g->dx[b][2];
        g->dr[b] = sqrt(sq);
        for (int c = 0; c < 3; c++) g->u[b][c] = g->dx[b][c] / g->dr[b];
    }
}

/* angle derivative block shared by intra_spcfw :156-173, intra_morse_harm :296-313 */
static void angle_terms(const geom_t* g, double* ang, double dth[3]) {
    double dr1sq = g->dx[0][0] * g->dx[0][0] + g->dx[0][1] * g->dx[0][1] + g->dx[0][2] * g->dx[0][2];
    double dr2sq = g->dx[1][0] * g->dx[1][0] + g->dx[1][1] * g->dx[1][1] + g->dx[1][2] * g->dx[1][2];
    double dr3sq = g->dx[2][0] * g->dx[2][0] + g->dx[2][1] * g->dx[2][1] + g->dx[2][2] * g->dx[2][2];
    double dr1a = g->dr[0], dr2a = g->dr[1], dr3a = g->dr[2];
    double u = (dr1sq + dr2sq - dr3sq);
    double vv = (2.0 * dr1a * dr2a);
    double v2 = 1.0 / (vv * vv);
    double arg = u / vv;
    double darg = -1.0 / sqrt(1.0 - arg * arg);
    *ang = acos(arg);
    double up = 2.0 * dr1a, vp = 2.0 * dr2a;
    double grad = (up * vv - vp * u) * v2;
    dth[0] = darg * grad;
    up = 2.0 * dr2a; vp = 2.0 * dr1a;
    grad = (up * vv - vp * u) * v2;
    dth[1] = darg * grad;
    up = -2.0 * dr3a;
    grad = (up * vv) * v2;
    dth[2] = darg * grad;
}

static void scatter_bonds(double* dvdr, int j, const double dv[3][3]) {
    for (int c = 0; c < 3; c++) {
        RX(dvdr, c, j) = RX(dvdr, c, j) - dv[0][c] - dv[1][c];
        RX(dvdr, c, j + 1) = RX(dvdr, c, j + 1) + dv[0][c] - dv[2][c];
        RX(dvdr, c, j + 2) = RX(dvdr, c, j + 2) + dv[1][c] + dv[2][c];
    }
}

static int intra_select(const rpmd_params* P) {
    if (P->intra_model != RPMD_INTRA_AUTO) return P->intra_model;
    if (P->alp == 0.0) return RPMD_INTRA_HARMONIC;            /* src/md_forces.f90:254-256 */
    return (P->alpb != 0.0) ? RPMD_INTRA_MORSE : RPMD_INTRA_MORSE_HARM; /* :258-262 */
}

static void intra(const rpmd_params* P, const double* r, double* dvdr, int na, double* vout, double* vir) {
    int model = intra_select(P);
    double v = 0.0;
    memset(vir, 0, sizeof(double) * 9);
    double theta = P->theta, reoh = P->reoh;
    for (int j = 0; j < na; j += 3) {
        geom_t g;
        mol_geom(r, j, &g);
        double dvb[3] = {0, 0, 0};
        if (model == RPMD_INTRA_HARMONIC) {
            /* intra_harmonic :23-95 */
            const double apot = 0.5993, bpot = 0.1466, cpot = -0.09435, dpot = 0.04984;
            double rehh = 2.0 * sin(0.5 * theta) * reoh;
            double dr1 = g.dr[0] - reoh, dr2 = g.dr[1] - reoh, dr3 = g.dr[2] - rehh;
            v = v + 0.5 * apot * (dr1 * dr1 + dr2 * dr2);
            v = v + 0.5 * bpot * dr3 * dr3;
            v = v + cpot * (dr1 + dr2) * dr3;
            v = v + dpot * dr1 * dr2;
            dvb[0] = apot * dr1 + cpot * dr3 + dpot * dr2;
            dvb[1] = apot * dr2 + cpot * dr3 + dpot * dr1;
            dvb[2] = bpot * dr3 + cpot * (dr1 + dr2);
        } else {
            double ang, dth[3];
            angle_terms(&g, &ang, dth);
            double dang = ang - theta;
            double dr1 = g.dr[0] - reoh, dr2 = g.dr[1] - reoh;
            if (model == RPMD_INTRA_SPCFW) {
                /* intra_spcfw :175-180 */
                v = v + 0.5 * P->apot * (dr1 * dr1 + dr2 * dr2);
                v = v + 0.5 * P->bpot * (dang * dang);
                dvb[0] = P->apot * dr1 + P->bpot * dang * dth[0];
                dvb[1] = P->apot * dr2 + P->bpot * dang * dth[1];
                dvb[2] = P->bpot * dang * dth[2];
            } else {
                /* intra_morse_harm :315-325 and intra_morse :315-325 */
                double de = P->apot, deb = P->bpot, alp = P->alp;
                double alp2 = alp * alp, alp3 = alp * alp2, alp4 = alp3 * alp;
                double f1 = 7.0 / 12.0, f2 = 7.0 / 3.0;
                double drasq = dr1 * dr1, drbsq = dr2 * dr2;
                v = v + de * (alp2 * drasq - alp3 * dr1 * drasq + f1 * alp4 * drasq * drasq);
                v = v + de * (alp2 * drbsq - alp3 * dr2 * drbsq + f1 * alp4 * drbsq * drbsq);
                double a1 = de * (2.0 * alp2 * dr1 - 3.0 * alp3 * drasq + f2 * alp4 * dr1 * drasq);
                double a2 = de * (2.0 * alp2 * dr2 - 3.0 * alp3 * drbsq + f2 * alp4 * dr2 * drbsq);
                double a3;
                if (model == RPMD_INTRA_MORSE_HARM) {
                    v = v + deb * (dang * dang);
                    a3 = 2.0 * deb * dang;
                } else {
                    double ab = P->alpb, ab2 = ab * ab, ab3 = ab2 * ab, ab4 = ab2 * ab2;
                    v = v + deb * (ab2 * dang * dang - ab3 * dang * dang * dang +
                                   f1 * ab4 * dang * dang * dang * dang);
                    a3 = deb * (2.0 * ab2 * dang - 3.0 * ab3 * dang * dang + f2 * ab4 * dang * dang * dang);
                }
                dvb[0] = a1 + a3 * dth[0];
                dvb[1] = a2 + a3 * dth[1];
                dvb[2] = a3 * dth[2];
            }
        }
        double dv[3][3];
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) dv[b][c] = dvb[b] * g.u[b][c];
        scatter_bonds(dvdr, j, dv);
        if (model == RPMD_INTRA_SPCFW) {
            /* W = sum r_i . dvdr_i, diagonal only (intra_spcfw :215-220) */
            for (int c = 0; c < 3; c++)
                VIR(vir, c, c) += RX(r, c, j) * (-dv[0][c] - dv[1][c]) +
                                  RX(r, c, j + 1) * (+dv[0][c] - dv[2][c]) +
                                  RX(r, c, j + 2) * (+dv[1][c] + dv[2][c]);
        } else {
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++)
                    VIR(vir, a, c) += g.dx[0][a] * dv[0][c] + g.dx[1][a] * dv[1][c] + g.dx[2][a] * dv[2][c];
        }
    }
    *vout = v;
}

/* ======================================================================================
 * TEM short-range Coulomb correction (src/correct_tem.f90)
 * ==================================================================================== */
typedef struct {
    const double *r, *z;
    double* dvdr;
    double v, vir[9];
    double rcutsq, a1, a2, a3;
    double box[3], onbox[3];
} tem_ctx;

/* pair body, dx = r_i - r_j, dvdr_i += dv (ewc_basic_tem :130-172) */
static void tem_pair(int i1, int j1, void* vctx) {
    tem_ctx* c = (tem_ctx*)vctx;
    int i = i1 - 1, j = j1 - 1;
    if (i / 3 == j / 3) return;
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = RX(c->r, k, i) - RX(c->r, k, j);
    mimage(c->box, c->onbox, d);
    double dx = d[0], dy = d[1], dz = d[2];
    double drsq = dx * dx + dy * dy + dz * dz;
    if (drsq < c->rcutsq) {
        double dr = sqrt(drsq), onr = 1.0 / dr, zij = c->z[i] * c->z[j];
        double vc = zij * onr, dvc = -vc * onr, dr3 = drsq * dr, dr4 = dr3 * dr;
        double f = 1.0 - c->a1 * dr + c->a2 * dr3 - c->a3 * dr4;
        double df = -c->a1 + 3.0 * c->a2 * drsq - 4.0 * c->a3 * dr3;
        double dv = vc * f, dvr = vc * df + dvc * f;
        double dvx = dvr * dx * onr, dvy = dvr * dy * onr, dvz = dvr * dz * onr;
        c->v = c->v + dv;
        RX(c->dvdr, 0, i) += dvx; RX(c->dvdr, 1, i) += dvy; RX(c->dvdr, 2, i) += dvz;
        RX(c->dvdr, 0, j) -= dvx; RX(c->dvdr, 1, j) -= dvy; RX(c->dvdr, 2, j) -= dvz;
        VIR(c->vir, 0, 0) += dx * dvx; VIR(c->vir, 1, 1) += dy * dvy; VIR(c->vir, 2, 2) += dz * dvz;
        VIR(c->vir, 0, 1) += dx * dvy; VIR(c->vir, 0, 2) += dx * dvz; VIR(c->vir, 1, 0) += dy * dvx;
        VIR(c->vir, 1, 2) += dy * dvz; VIR(c->vir, 2, 0) += dz * dvx; VIR(c->vir, 2, 1) += dz * dvy;
    }
}

/* ewc_driver_tem src/correct_tem.f90:1-71 (r is a private copy; M-site substituted here) */
static int ewc_driver_tem(const rpmd_params* P, double* r, const double* z, double* dvdr, double* v,
                          double* vir, int na, const double box[3], const nlist_t* L) {
    if (P->wm > 0.0 || P->wh > 0.0) return RPMD_ERR_SMEARED;
    double alpha = P->alpha, alpha2 = 0.5 * (1.0 - alpha);
    int msite = fabs(alpha - 1.0) > 1e-4;
    double* ro = NULL;
    if (msite) {
        ro = (double*)malloc(sizeof(double) * na);
        for (int i = 0; i < na; i += 3)
            for (int c = 0; c < 3; c++) {
                ro[i + c] = RX(r, c, i);
                RX(r, c, i) = alpha * RX(r, c, i) + alpha2 * (RX(r, c, i + 1) + RX(r, c, i + 2));
            }
    }
    tem_ctx c;
    memset(&c, 0, sizeof c);
    c.r = r; c.z = z; c.dvdr = dvdr;
    double sig = P->sig;
    c.rcutsq = sig * sig;
    c.a1 = 2.0 / sig;
    c.a2 = 2.0 / (sig * sig * sig);
    c.a3 = 1.0 / (sig * sig * sig * sig);
    for (int d = 0; d < 3; d++) { c.box[d] = box[d]; c.onbox[d] = 1.0 / box[d]; }
    double bmax = box[0];
    if (box[1] > bmax) bmax = box[1];
    if (box[2] > bmax) bmax = box[2];
    int err = 0;
    memset(dvdr, 0, sizeof(double) * 3 * na);
    if (L && L->list[1] != 0) {
        /* ewc_list_tem :179-296: rcut check uses the largest side */
        if (sig > 0.5 * bmax) err = RPMD_ERR_EWC_CUTOFF;
        for (int j = 1; j <= na && !err; j++) {
            int ibeg = L->point[j], iend = L->point[j + 1] - 1;
            if (ibeg <= iend) {
                double dj[3] = {RX(dvdr, 0, j - 1), RX(dvdr, 1, j - 1), RX(dvdr, 2, j - 1)};
                for (int inab = ibeg; inab <= iend; inab++) {
                    int i = L->list[inab];
                    if ((i - 1) / 3 == (j - 1) / 3) continue;
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = RX(r, k, i - 1) - RX(r, k, j - 1);
                    mimage(box, c.onbox, d);
                    double dx = d[0], dy = d[1], dz = d[2];
                    double drsq = dx * dx + dy * dy + dz * dz;
                    if (drsq < c.rcutsq) {
                        double dr = sqrt(drsq), onr = 1.0 / dr, zij = z[i - 1] * z[j - 1];
                        double vc = zij * onr, dvc = -vc * onr, dr3 = drsq * dr, dr4 = dr3 * dr;
                        double f = 1.0 - c.a1 * dr + c.a2 * dr3 - c.a3 * dr4;
                        double df = -c.a1 + 3.0 * c.a2 * drsq - 4.0 * c.a3 * dr3;
                        double dv = vc * f, dvr = vc * df + dvc * f;
                        double dvx = dvr * dx * onr, dvy = dvr * dy * onr, dvz = dvr * dz * onr;
                        c.v = c.v + dv;
                        RX(dvdr, 0, i - 1) += dvx; RX(dvdr, 1, i - 1) += dvy; RX(dvdr, 2, i - 1) += dvz;
                        dj[0] -= dvx; dj[1] -= dvy; dj[2] -= dvz;
                        VIR(c.vir, 0, 0) += dx * dvx; VIR(c.vir, 1, 1) += dy * dvy; VIR(c.vir, 2, 2) += dz * dvz;
                        VIR(c.vir, 0, 1) += dx * dvy; VIR(c.vir, 0, 2) += dx * dvz; VIR(c.vir, 1, 0) += dy * dvx;
                        VIR(c.vir, 1, 2) += dy * dvz; VIR(c.vir, 2, 0) += dz * dvx; VIR(c.vir, 2, 1) += dz * dvy;
                    }
                }
                RX(dvdr, 0, j - 1) = dj[0]; RX(dvdr, 1, j - 1) = dj[1]; RX(dvdr, 2, j - 1) = dj[2];
            }
        }
    } else if (3.0 * sig > bmax) {
        /* ewc_basic_tem :73-177: the 'boxmin' there is the largest side */
        if (sig > 0.5 * bmax) err = RPMD_ERR_EWC_CUTOFF;
        for (int j = 1; j <= na && !err; j++)
            for (int i = 1; i <= j - 1; i++) tem_pair(i, j, &c);
    } else {
        /* ewc_cell_tem :299-498 */
        cells_t C;
        memset(&C, 0, sizeof C);
        cell_setup(box, sig, C.mc, C.nc, &C.nn);
        cell_list(r, na, box, 1, &C);
        cells_pairs_percell(&C, tem_pair, &c);
        cells_free(&C);
    }
    if (msite) {
        for (int j = 0; j < na; j += 3)
            for (int k = 0; k < 3; k++) {
                RX(dvdr, k, j + 1) = RX(dvdr, k, j + 1) + alpha2 * RX(dvdr, k, j);
                RX(dvdr, k, j + 2) = RX(dvdr, k, j + 2) + alpha2 * RX(dvdr, k, j);
                RX(dvdr, k, j) = alpha * RX(dvdr, k, j);
                RX(r, k, j) = ro[j + k];
            }
        free(ro);
    }
    *v = c.v;
    memcpy(vir, c.vir, sizeof(double) * 9);
    return err;
}

/* ======================================================================================
 * potenl_opt src/md_forces.f90:113-302 and forces :1-110
 * ==================================================================================== */
static int potenl_opt(const rpmd_params* P, const double* rin, double* v, double* dvdr, double* vir,
                      int na, int nb, const double box[3], const double* z, const nlist_t* L, int iopt) {
    (void)nb;
    *v = 0.0;
    memset(dvdr, 0, sizeof(double) * 3 * na);
    memset(vir, 0, sizeof(double) * 9);
    int nm = na / 3;
    if (3 * nm != na) return RPMD_ERR_NA_NOT_MULT3;
    if (iopt <= 2) {
        double alpha = P->alpha, alpha2 = 0.5 * (1.0 - alpha);
        int* mol = (int*)malloc(sizeof(int) * na);
        for (int j = 0; j < na; j++) mol[j] = j / 3 + 1;
        double* r = (double*)malloc(sizeof(double) * 3 * na);
        memcpy(r, rin, sizeof(double) * 3 * na);
        for (int i = 0; i < na; i += 3)
            for (int c = 0; c < 3; c++)
                RX(r, c, i) = alpha * RX(rin, c, i) + alpha2 * (RX(rin, c, i + 1) + RX(rin, c, i + 2));
        double vew, vir_ew[9];
        if (P->iamcub) ewald_cubic(P, na, mol, z, r, &vew, vir_ew, dvdr, box[0]);
        else ewald_nc(P, na, mol, z, r, &vew, vir_ew, dvdr, box);
        *v = *v + vew;
        for (int k = 0; k < 9; k++) vir[k] = vir[k] + vir_ew[k];
        for (int j = 0; j < na; j += 3)
            for (int i = 0; i < 3; i++) {
                RX(dvdr, i, j + 1) = RX(dvdr, i, j + 1) + alpha2 * RX(dvdr, i, j);
                RX(dvdr, i, j + 2) = RX(dvdr, i, j + 2) + alpha2 * RX(dvdr, i, j);
                RX(dvdr, i, j) = alpha * RX(dvdr, i, j);
            }
        free(r);
        free(mol);
    }
    if (iopt == 0 || iopt == 1 || iopt == 3) {
        double* dvdroo = (double*)malloc(sizeof(double) * 3 * na);
        double voo, vir_oo[9];
        oo_driver(P, rin, dvdroo, &voo, vir_oo, L, na, box, 3);
        for (int k = 0; k < 3 * na; k++) dvdr[k] = dvdr[k] + dvdroo[k];
        *v = *v + voo;
        for (int k = 0; k < 9; k++) vir[k] = vir[k] + vir_oo[k];
        free(dvdroo);
    }
    if (iopt == 0 || iopt == 4) {
        double vint, vir_int[9];
        intra(P, rin, dvdr, na, &vint, vir_int);
        *v = *v + vint;
        for (int k = 0; k < 9; k++) vir[k] = vir[k] + vir_int[k];
    }
    return 0;
}

int orc_forces(const rpmd_params* P, const double* r, int nb, int na, const double box[3],
               const double* z, int iopt, double* vout, double* dvdr, double* virial) {
    if (iopt < 0 || iopt > 5) return RPMD_ERR_IOPT;
    if (na % 3 != 0) return RPMD_ERR_NA_NOT_MULT3;
    double rgmax = 5.0, v = 0.0, vir[9];
    memset(dvdr, 0, sizeof(double) * 3 * na * nb);
    memset(virial, 0, sizeof(double) * 9);
    nlist_t L;
    nlist_alloc(&L, na);
    int err = 0;
    double boxmin = box[0];
    if (box[1] < boxmin) boxmin = box[1];
    if (box[2] < boxmin) boxmin = box[2];
    if (iopt == 3 || iopt <= 1) {
        if (nb > 3) {
            double* rc = (double*)calloc((size_t)3 * na, sizeof(double));
            for (int j = 0; j < nb; j++)
                for (int k = 0; k < 3 * na; k++) rc[k] = rc[k] + r[(size_t)j * 3 * na + k];
            for (int k = 0; k < 3 * na; k++) rc[k] = rc[k] / (double)nb;
            double cut = P->rcut;
            if (0.5 * boxmin < cut) cut = 0.5 * boxmin;
            err = nei_driver(rc, na, &L, 3, cut, box);
            free(rc);
        } else {
            L.list[1] = 0;
        }
    }
    if (!err && iopt != 5) {
        for (int k = 0; k < nb && !err; k++) {
            double dv;
            err = potenl_opt(P, r + (size_t)k * 3 * na, &dv, dvdr + (size_t)k * 3 * na, vir, na, nb,
                             box, z, (iopt == 3 || iopt <= 1) ? &L : NULL, iopt);
            v = v + dv;
            for (int m = 0; m < 9; m++) virial[m] = virial[m] + vir[m];
        }
    } else if (!err) {
        if (nb > 1) {
            double* rc = (double*)calloc((size_t)3 * na, sizeof(double));
            for (int j = 0; j < nb; j++)
                for (int k = 0; k < 3 * na; k++) rc[k] = rc[k] + r[(size_t)j * 3 * na + k];
            for (int k = 0; k < 3 * na; k++) rc[k] = rc[k] / (double)nb;
            double cut = P->sig + rgmax;
            if (cut > 0.5 * boxmin) cut = 0.5 * boxmin; /* 'ewc cutoff is too big for box' */
            err = nei_driver(rc, na, &L, 1, cut, box);
            free(rc);
        } else {
            L.list[1] = 0;
        }
        double* rb = (double*)malloc(sizeof(double) * 3 * na);
        for (int k = 0; k < nb && !err; k++) {
            double dv;
            memcpy(rb, r + (size_t)k * 3 * na, sizeof(double) * 3 * na);
            err = ewc_driver_tem(P, rb, z, dvdr + (size_t)k * 3 * na, &dv, vir, na, box,
                                 nb > 1 ? &L : NULL);
            v = v + dv;
            for (int m = 0; m < 9; m++) virial[m] = virial[m] + vir[m];
        }
        free(rb);
    }
    nlist_free(&L);
    *vout = v / (double)nb;
    for (int m = 0; m < 9; m++) virial[m] = virial[m] / (double)nb;
    return err;
}

/* ======================================================================================
 * realft src/fourier.f90:16-63: FFTW R2HC (mode +1) / HC2R (mode -1), times 1/sqrt(n),
 * restated from FFTW's documented definition as a direct DFT in long double.
 * ==================================================================================== */
int orc_realft(double* data, int m, int n, int mode) {
    if (n > 256) return RPMD_ERR_REALFT_N;
    if (mode != 1 && mode != -1) return RPMD_ERR_REALFT_MODE;
    long double pi = acosl(-1.0L);
    long double* cs = (long double*)malloc(sizeof(long double) * n);
    long double* sn = (long double*)malloc(sizeof(long double) * n);
    for (int t = 0; t < n; t++) {
        cs[t] = cosl(2.0L * pi * t / n);
        sn[t] = sinl(2.0L * pi * t / n);
    }
    double* x = (double*)malloc(sizeof(double) * n);
    long double* y = (long double*)malloc(sizeof(long double) * n);
    double scale = sqrt(1.0 / (double)n);
    for (int k = 0; k < m; k++) {
        for (int j = 0; j < n; j++) x[j] = data[(size_t)j * m + k];
        if (mode == 1) {
            for (int q = 0; q <= n / 2; q++) {
                long double s = 0;
                for (int j = 0; j < n; j++) s += x[j] * cs[(long long)j * q % n];
                y[q] = s;
            }
            for (int q = 1; q <= (n - 1) / 2; q++) {
                long double s = 0;
                for (int j = 0; j < n; j++) s -= x[j] * sn[(long long)j * q % n];
                y[n - q] = s;
            }
        } else {
            for (int j = 0; j < n; j++) {
                long double s = x[0];
                for (int q = 1; q <= (n - 1) / 2; q++)
                    s += 2.0L * (x[q] * cs[(long long)j * q % n] - x[n - q] * sn[(long long)j * q % n]);
                if (n % 2 == 0) s += (j % 2 == 0 ? 1.0L : -1.0L) * x[n / 2];
                y[j] = s;
            }
        }
        for (int j = 0; j < n; j++) data[(size_t)j * m + k] = scale * (double)y[j];
    }
    free(cs); free(sn); free(x); free(y);
    return 0;
}

/* ring_rpmd src/md_pathi.f90:190-231: monod(3,4,nb) column-major */
void orc_ring_rpmd(const double* mass, double* monod, int na, int nb, double dt, double beta) {
    (void)na;
    const double hbar = 1.0;
    double pi = acos(-1.0);
#define MON(j, a, k) monod[((k) * 4 + (a)) * 3 + (j)]
    memset(monod, 0, sizeof(double) * 12 * nb);
    for (int j = 0; j < 3; j++) {
        MON(j, 0, 0) = 1.0;
        MON(j, 1, 0) = 0.0;
        MON(j, 2, 0) = dt / mass[j];
        MON(j, 3, 0) = 1.0;
        if (nb > 1) {
            double betan = beta / (double)nb, twown = 2.0 / (betan * hbar), pibyn = pi / (double)nb;
            for (int k = 1; k <= nb / 2; k++) {
                double wk = twown * sin(k * pibyn), wt = wk * dt, wm = wk * mass[j];
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
#undef MON
}

/* ======================================================================================
 * RP contraction: src/evolve_pi.f90:324-431
 * ==================================================================================== */
int orc_rp_contract_nm(const rpmd_params* P, const double* rnm, double* v, double* dvdr, int nb, int na,
                       const double box[3], const double* z, double* vir, int nbdf, int iopt) {
    if ((nbdf - 1) % 2 != 0) return RPMD_ERR_NBDF_EVEN;
    if (nbdf > nb) return RPMD_ERR_NBDF_TOO_LARGE;
    size_t A = (size_t)3 * na;
    double* rb = (double*)calloc(A * nbdf, sizeof(double));
    double* dvdrb = (double*)calloc(A * nbdf, sizeof(double));
    memset(dvdr, 0, sizeof(double) * A * nb);
    /* ring_contract :405-431 */
    memcpy(rb, rnm, sizeof(double) * A);
    for (int k = 1; k <= (nbdf - 1) / 2; k++) {
        memcpy(rb + A * k, rnm + A * k, sizeof(double) * A);
        memcpy(rb + A * (nbdf - k), rnm + A * (nb - k), sizeof(double) * A);
    }
    orc_realft(rb, (int)A, nbdf, -1);
    double sf = sqrt((double)nbdf / (double)nb);
    for (size_t i = 0; i < A * nbdf; i++) rb[i] = sf * rb[i];
    int err = orc_forces(P, rb, nbdf, na, box, z, iopt, v, dvdrb, vir);
    /* force_contract :377-403 */
    orc_realft(dvdrb, (int)A, nbdf, +1);
    sf = sqrt((double)nb / (double)nbdf);
    for (size_t i = 0; i < A; i++) dvdr[i] = sf * dvdrb[i];
    for (int k = 1; k <= (nbdf - 1) / 2; k++)
        for (size_t i = 0; i < A; i++) {
            dvdr[A * k + i] = sf * dvdrb[A * k + i];
            dvdr[A * (nb - k) + i] = sf * dvdrb[A * (nbdf - k) + i];
        }
    free(rb);
    free(dvdrb);
    return err;
}

/* full_forces src/md_forces.f90:305-472 (rpmddft = 0 branch) */
int orc_full_forces(const rpmd_params* P, const double* rin, int na, int nb, double* v, double* vew,
                    double* voo, double* vint, double* vir, const double* z, const double box[3],
                    double* dvdr, double* dvdr2) {
    size_t A = (size_t)3 * na, N = A * nb;
    double* r = (double*)malloc(sizeof(double) * N);
    memcpy(r, rin, sizeof(double) * N);
    double* dvdre = (double*)calloc(N, sizeof(double));
    double* dvdrl = (double*)calloc(N, sizeof(double));
    double vir_itr[9] = {0}, vir_oo[9] = {0}, vir_ew[9] = {0}, vir_ewc[9] = {0};
    double ve = 0.0;
    *vew = 0.0; *voo = 0.0; *vint = 0.0;
    memset(dvdr, 0, sizeof(double) * N);
    int err = orc_forces(P, r, nb, na, box, z, 4, vint, dvdr2, vir_itr);
    int nbdf1 = P->nbdf1, nbdf2 = P->nbdf2;
    if (!err && nb > 1) {
        if (nbdf1 > 0 || nbdf2 > 0) {
            orc_realft(r, (int)A, nb, +1);
            if (nbdf1 > 0) {
                err = orc_rp_contract_nm(P, r, vew, dvdr, nb, na, box, z, vir_ew, nbdf1, 2);
                if (!err && P->sig > 0.0) {
                    err = orc_rp_contract_nm(P, r, &ve, dvdre, nb, na, box, z, vir_ewc, nbdf1, 5);
                    *vew = *vew - ve;
                    for (int m = 0; m < 9; m++) vir_ew[m] = vir_ew[m] - vir_ewc[m];
                    for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] - dvdre[i];
                }
            }
            if (!err && nbdf2 > 0)
                err = orc_rp_contract_nm(P, r, voo, dvdrl, nb, na, box, z, vir_oo, nbdf2, 3);
            orc_realft(r, (int)A, nb, -1);
        }
        if (nbdf1 > 0 && nbdf2 > 0) {
            for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdrl[i];
            orc_realft(dvdr, (int)A, nb, -1);
        } else if (nbdf1 > 0) {
            orc_realft(dvdr, (int)A, nb, -1);
        } else if (nbdf2 > 0) {
            orc_realft(dvdrl, (int)A, nb, -1);
        }
    }
    if (!err) {
        if (nb == 1 || nbdf1 == 0) {
            err = orc_forces(P, r, nb, na, box, z, 2, vew, dvdr, vir_ew);
        } else if (P->sig > 0.0) {
            err = orc_forces(P, r, nb, na, box, z, 5, &ve, dvdre, vir_ewc);
            *vew = *vew + ve;
            for (int m = 0; m < 9; m++) vir_ew[m] = vir_ew[m] + vir_ewc[m];
            for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdre[i];
        }
    }
    if (!err && (nb == 1 || nbdf2 == 0)) err = orc_forces(P, r, nb, na, box, z, 3, voo, dvdrl, vir_oo);
    if (nb == 1 || nbdf1 <= 0 || nbdf2 <= 0)
        for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdrl[i];
    *v = *vew + *voo + *vint;
    for (int m = 0; m < 9; m++) vir[m] = vir_oo[m] + vir_ew[m] + vir_itr[m];
    free(r); free(dvdre); free(dvdrl);
    return err;
}

/* ======================================================================================
 * RNG: src/random.f90 (ran2 = Numerical Recipes L'Ecuyer + Bays-Durham; polar Box-Muller)
 * ==================================================================================== */
void orc_rng_init(orc_rng* s) {
    memset(s, 0, sizeof *s);
    s->idum2 = 123456789;
}

double orc_ran2(orc_rng* s, double rmin, double rmax) {
    const int im1 = 2147483563, im2 = 2147483399, imm1 = im1 - 1, ia1 = 40014, ia2 = 40692,
              iq1 = 53668, iq2 = 52774, ir1 = 12211, ir2 = 3791, ntab = 32, ndiv = 1 + imm1 / ntab;
    const double am = 1.0 / im1, eps = 1.2e-7, rnmx = 1.0 - eps;
    int k, j;
    if (s->idum <= 0) {
        s->idum = (-s->idum > 1) ? -s->idum : 1;
        s->idum2 = s->idum;
        for (j = ntab + 8; j >= 1; j--) {
            k = s->idum / iq1;
            s->idum = ia1 * (s->idum - k * iq1) - k * ir1;
            if (s->idum < 0) s->idum += im1;
            if (j <= ntab) s->iv[j - 1] = s->idum;
        }
        s->iy = s->iv[0];
    }
    k = s->idum / iq1;
    s->idum = ia1 * (s->idum - k * iq1) - k * ir1;
    if (s->idum < 0) s->idum += im1;
    k = s->idum2 / iq2;
    s->idum2 = ia2 * (s->idum2 - k * iq2) - k * ir2;
    if (s->idum2 < 0) s->idum2 += im2;
    j = 1 + s->iy / ndiv;
    s->iy = s->iv[j - 1] - s->idum2;
    s->iv[j - 1] = s->idum;
    if (s->iy < 1) s->iy += imm1;
    double u = am * s->iy;
    return ((rmax - rmin) * (u < rnmx ? u : rnmx)) + rmin;
}

double orc_gaussian(orc_rng* s, double sigma) {
    if (s->iset == 0) {
        double v1, v2, rsq;
        do {
            v1 = 2.0 * orc_ran2(s, 0.0, 1.0) - 1.0;
            v2 = 2.0 * orc_ran2(s, 0.0, 1.0) - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (rsq >= 1.0 || rsq == 0.0);
        double fac = sqrt(-2.0 * log(rsq) / rsq);
        s->gset = v1 * fac;
        s->iset = 1;
        return sigma * v2 * fac;
    }
    s->iset = 0;
    return sigma * s->gset;
}

/* main.f90:298-299: irun = -irun; test = gaussian(irun, 1) */
void orc_rng_seed(orc_rng* s, int irun) {
    orc_rng_init(s);
    s->idum = -irun;
    (void)orc_gaussian(s, 1.0);
}

void orc_gaussian_fill(orc_rng* s, double* out, long long n) {
    for (long long i = 0; i < n; i++) out[i] = orc_gaussian(s, 1.0);
}

/* ======================================================================================
 * Thermostats
 * ==================================================================================== */
static void com_remove_centroid(double* p, const double* mass, int na) {
    double pcom[3] = {0, 0, 0}, mm = 0.0;
    for (int j = 0; j < na; j++) {
        for (int i = 0; i < 3; i++) pcom[i] = pcom[i] + RX(p, i, j);
        mm = mm + mass[j];
    }
    for (int i = 0; i < 3; i++) pcom[i] = pcom[i] * (1.0 / mm);
    for (int j = 0; j < na; j++) {
        mm = mass[j];
        for (int i = 0; i < 3; i++) RX(p, i, j) = RX(p, i, j) - pcom[i] * mm;
    }
}

/* therm_pro src/module_gle.f90:8-73 (PILE-L in the NM representation) */
void orc_therm_pro(double* p, const double* mass, double beta, double ctau, double dt, int na, int nb,
                   orc_rng* s) {
    const double hbar = 1.0;
    double betan = beta / nb, pi = acos(-1.0), twown = 2.0 / (betan * hbar), pibyn = pi / nb;
    double radtwo = sqrt(2.0), ck[3];
    for (int j = 0; j < 3; j++) ck[j] = sqrt(mass[j]);
    for (int k = 1; k <= nb; k++) {
        double wk = twown * sin((k - 1) * pibyn), gamma;
        if (k == 1) gamma = 2.0 * pi / ctau;
        else gamma = 2.0 * wk;
        double c1 = exp(-gamma * dt);
        double c2 = sqrt((1.0 - c1 * c1) * 0.5 / betan);
        if (k == 1 || ((nb % 2 == 0) && k == (nb / 2 + 1))) c2 = radtwo * c2;
        double* pk = p + (size_t)(k - 1) * 3 * na;
        for (int j = 0; j < na; j++) {
            double c3 = c2 * ck[j % 3];
            for (int i = 0; i < 3; i++) RX(pk, i, j) = c1 * RX(pk, i, j) + c3 * orc_gaussian(s, 1.0);
        }
    }
    com_remove_centroid(p, mass, na);
}

/* parinello_therm src/module_thermostat.f90:6-54 (Bussi global) */
void orc_parinello_therm(double* p, const double* mass, double tau, int na, int nb, double dt, orc_rng* s,
                         double beta, int nctot) {
    double betan = beta / (double)nb;
    int ndof = (int)(3 * na * (double)nb - nctot - 3);
    (void)orc_gaussian(s, 1.0); /* rt, drawn and unused */
    double ktav = (double)ndof / (2.0 * betan), snfm1 = 0.0;
    for (int i = 1; i <= ndof - 1; i++) {
        double g = orc_gaussian(s, 1.0);
        snfm1 = snfm1 + g * g;
    }
    double kt = 0.0;
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++)
            for (int i = 0; i < 3; i++) {
                double pv = p[((size_t)k * na + j) * 3 + i];
                kt = kt + pv * pv / (2.0 * mass[j]);
            }
    double ranno = orc_gaussian(s, 1.0), ransq = ranno * ranno;
    double c = exp(-dt / tau);
    double fac = (1.0 - c) * ktav / ((double)ndof * kt);
    double t1 = fac * (snfm1 + ransq);
    double t2 = 2 * ranno * sqrt(c * fac);
    double alpha = sqrt(c + t1 + t2);
    double plusmin = ranno + sqrt(c / fac);
    alpha = (plusmin >= 0.0) ? fabs(alpha) : -fabs(alpha);
    for (size_t i = 0; i < (size_t)3 * na * nb; i++) p[i] = alpha * p[i];
}

/* therm_pro_global src/module_gle.f90:75-143 (PILE-G) */
void orc_therm_pro_global(double* p, const double* mass, double beta, double ctau, double dt, int na, int nb,
                          orc_rng* s) {
    const double hbar = 1.0;
    double betan = beta / nb, pi = acos(-1.0), twown = 2.0 / (betan * hbar), pibyn = pi / nb;
    double radtwo = sqrt(2.0), ck[3];
    for (int j = 0; j < 3; j++) ck[j] = sqrt(mass[j]);
    for (int k = 2; k <= nb; k++) {
        double wk = twown * sin((k - 1) * pibyn), gamma = 2.0 * wk;
        double c1 = exp(-gamma * dt);
        double c2 = sqrt((1.0 - c1 * c1) * 0.5 / betan);
        if ((nb % 2 == 0) && k == (nb / 2 + 1)) c2 = radtwo * c2;
        double* pk = p + (size_t)(k - 1) * 3 * na;
        for (int j = 0; j < na; j++) {
            double c3 = c2 * ck[j % 3];
            for (int i = 0; i < 3; i++) RX(pk, i, j) = c1 * RX(pk, i, j) + c3 * orc_gaussian(s, 1.0);
        }
    }
    orc_parinello_therm(p, mass, ctau, na, 1, dt, s, betan, 0);
    com_remove_centroid(p, mass, na);
}

/* parinello_therm_loc src/module_thermostat.f90:56-82 */
void orc_parinello_therm_loc(double* p, const double* mass, double tau, int na, int nb, double dt,
                             orc_rng* s, double beta) {
    double gamma = 1.0 / (2.0 * tau), betan = beta / (double)nb;
    double c1 = exp(-0.5 * gamma * dt), c2 = sqrt(1.0 - c1 * c1);
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++) {
            double c3 = c2 * sqrt(mass[j] / betan);
            for (int i = 0; i < 3; i++) {
                double* pv = &p[((size_t)k * na + j) * 3 + i];
                *pv = c1 * *pv + c3 * orc_gaussian(s, 1.0);
            }
        }
}

/* pshift src/module_thermostat.f90:229-264 */
void orc_pshift(double* p, const double* mass, int na, int nb) {
    double vx = 0, vy = 0, vz = 0, totmass = 0;
    for (int j = 0; j < na; j++) totmass = totmass + mass[j];
    totmass = (double)nb * totmass;
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++) {
            double* pv = &p[((size_t)k * na + j) * 3];
            vx = vx + pv[0]; vy = vy + pv[1]; vz = vz + pv[2];
        }
    vx = vx / totmass; vy = vy / totmass; vz = vz / totmass;
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++) {
            double* pv = &p[((size_t)k * na + j) * 3];
            pv[0] = pv[0] - mass[j] * vx;
            pv[1] = pv[1] - mass[j] * vy;
            pv[2] = pv[2] - mass[j] * vz;
        }
}

/* sample_rpmd src/module_thermostat.f90:199-226 */
void orc_sample_rpmd(double* p, int na, int nb, const double* mass, double beta, orc_rng* s) {
    double betan = beta / (double)nb;
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++)
            for (int l = 0; l < 3; l++) {
                double sigma = sqrt(mass[j] / betan);
                p[((size_t)k * na + j) * 3 + l] = orc_gaussian(s, sigma);
            }
    orc_pshift(p, mass, na, nb);
}

/* matrix_exp src/module_gle.f90:323-354 (Taylor j terms, k squarings) */
void orc_matrix_exp(const double* M, int n, int j, int k, double* EM) {
    double tc[64], SM[64], T[64];
    tc[0] = 1.0;
    for (int i = 1; i <= j; i++) tc[i] = tc[i - 1] / (double)i;
    double sc = (double)(1.0f / powf(2.0f, (float)k));
    for (int i = 0; i < n * n; i++) SM[i] = M[i] * sc;
    memset(EM, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) EM[i * n + i] = tc[j];
    for (int p = j; p >= 1; p--) {
        for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++) {
                double s = 0;
                for (int c = 0; c < n; c++) s += SM[c * n + a] * EM[b * n + c];
                T[b * n + a] = s;
            }
        memcpy(EM, T, sizeof(double) * n * n);
        for (int i = 0; i < n; i++) EM[i * n + i] = EM[i * n + i] + tc[p - 1];
    }
    for (int p = 1; p <= k; p++) {
        for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++) {
                double s = 0;
                for (int c = 0; c < n; c++) s += EM[c * n + a] * EM[b * n + c];
                T[b * n + a] = s;
            }
        memcpy(EM, T, sizeof(double) * n * n);
    }
}

/* cholesky src/module_gle.f90:357-378 (column-major n x n) */
void orc_cholesky(const double* SST, double* S, int n) {
#define MI(a, i, j) (a)[(j) * n + (i)]
    memset(S, 0, sizeof(double) * n * n);
    MI(S, 0, 0) = sqrt(MI(SST, 0, 0));
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < i; j++) {
            MI(S, i, j) = MI(SST, i, j);
            for (int k = 0; k < j; k++) MI(S, i, j) = MI(S, i, j) - MI(S, i, k) * MI(S, j, k);
            MI(S, i, j) = MI(S, i, j) / MI(S, j, j);
        }
        MI(S, i, i) = MI(SST, i, i);
        for (int k = 0; k < i; k++) MI(S, i, i) = MI(S, i, i) - MI(S, i, k) * MI(S, i, k);
        MI(S, i, i) = sqrt(MI(S, i, i));
    }
#undef MI
}

/* therm_gle_init src/module_gle.f90:146-227; A is (ns+1)^2 column-major as read from GLE-A,
 * C = (nb/beta) I (no GLE-C file). dt is the half step passed by main.f90:609. */
int orc_gle_init(orc_gle* g, const double* A, int ns, double ttau, int na, int nb, double dt, orc_rng* s,
                 double beta) {
    int n = ns + 1;
    if (n > 8) return RPMD_ERR_ARG;
    g->ns = ns; g->na = na; g->nb = nb;
    double gA[64], gC[64], M[64], tmp[64], tmp2[64];
    double fac = (double)(2.0f * 3.1415927f) / ttau; /* single-precision literals (:173) */
    for (int i = 0; i < n * n; i++) gA[i] = A[i] * fac;
    memset(gC, 0, sizeof gC);
    for (int i = 0; i < n; i++) gC[i * n + i] = ((double)nb / beta);
    for (int i = 0; i < n * n; i++) M[i] = -dt * gA[i];
    orc_matrix_exp(M, n, 15, 15, g->gT);
    /* gA = gC - gT gC gT^T */
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            double sm = 0;
            for (int c = 0; c < n; c++) sm += gC[c * n + a] * g->gT[c * n + b]; /* (gC gT^T)(a,b) */
            tmp[b * n + a] = sm;
        }
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            double sm = 0;
            for (int c = 0; c < n; c++) sm += g->gT[c * n + a] * tmp[b * n + c];
            tmp2[b * n + a] = sm;
        }
    for (int i = 0; i < n * n; i++) gA[i] = gC[i] - tmp2[i];
    orc_cholesky(gA, g->gS, n);
    double Cc[64];
    orc_cholesky(gC, Cc, n);
    g->gp = (double*)calloc((size_t)3 * na * nb * n, sizeof(double));
    size_t NN = (size_t)3 * na * nb;
    double gr[8];
    for (int h = 0; h < 3; h++)
        for (int k = 0; k < na; k++)
            for (int j = 0; j < nb; j++) {
                for (int i = 0; i < n; i++) gr[i] = orc_gaussian(s, 1.0);
                for (int i = 0; i < n; i++) {
                    double sm = 0;
                    for (int c = 0; c < n; c++) sm += Cc[c * n + i] * gr[c];
                    g->gp[(size_t)i * NN + ((size_t)j * na + k) * 3 + h] = sm;
                }
            }
    return 0;
}

void orc_gle_free(orc_gle* g) { free(g->gp); g->gp = NULL; }

/* pshift_gle src/module_gle.f90:285-320 */
static void pshift_gle(double* gp, const double* mass, int na, int nb, int ns) {
    size_t NN = (size_t)3 * na * nb;
    double mm = 0.0;
    for (int i = 0; i < na; i++) mm = mm + mass[i];
    mm = 1.0 / sqrt(nb * mm);
    for (int i = 0; i <= ns; i++) {
        double pcom[3] = {0, 0, 0};
        double* g = gp + (size_t)i * NN;
        for (int j = 0; j < na; j++) {
            double mfac = sqrt(mass[j]);
            for (int k = 0; k < nb; k++)
                for (int h = 0; h < 3; h++) pcom[h] = pcom[h] + g[((size_t)k * na + j) * 3 + h] * mfac;
        }
        for (int h = 0; h < 3; h++) pcom[h] = pcom[h] * mm * mm;
        for (int j = 0; j < na; j++) {
            double mfac = sqrt(mass[j]);
            for (int k = 0; k < nb; k++)
                for (int h = 0; h < 3; h++) g[((size_t)k * na + j) * 3 + h] -= pcom[h] * mfac;
        }
    }
}

/* therm_gle src/module_gle.f90:229-283 */
void orc_therm_gle(orc_gle* g, double* p, const double* mass, int na, int nb, orc_rng* s) {
    int n1 = g->ns + 1;
    size_t NN = (size_t)3 * na * nb;
    for (int k = 0; k < na; k++) {
        double mfac = 1.0 / sqrt(mass[k]);
        for (int h = 0; h < 3; h++)
            for (int j = 0; j < nb; j++) {
                size_t o = ((size_t)j * na + k) * 3 + h;
                g->gp[o] = p[o] * mfac;
            }
    }
    double* ngp = (double*)malloc(sizeof(double) * NN * n1);
    for (size_t e = 0; e < NN; e++)
        for (int a = 0; a < n1; a++) {
            double sm = 0;
            for (int c = 0; c < n1; c++) sm += g->gp[(size_t)c * NN + e] * g->gT[c * n1 + a];
            ngp[(size_t)a * NN + e] = sm;
        }
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < nb; j++)
            for (int k = 0; k < na; k++)
                for (int h = 0; h < 3; h++) g->gp[(size_t)i * NN + ((size_t)j * na + k) * 3 + h] = orc_gaussian(s, 1.0);
    for (size_t e = 0; e < NN; e++)
        for (int a = 0; a < n1; a++) {
            double sm = 0;
            for (int c = 0; c < n1; c++) sm += g->gp[(size_t)c * NN + e] * g->gS[c * n1 + a];
            ngp[(size_t)a * NN + e] = ngp[(size_t)a * NN + e] + sm;
        }
    memcpy(g->gp, ngp, sizeof(double) * NN * n1);
    free(ngp);
    pshift_gle(g->gp, mass, na, nb, g->ns);
    for (int k = 0; k < na; k++) {
        double mfac = sqrt(mass[k]);
        for (int h = 0; h < 3; h++)
            for (int j = 0; j < nb; j++) {
                size_t o = ((size_t)j * na + k) * 3 + h;
                p[o] = g->gp[o] * mfac;
            }
    }
}

/* ======================================================================================
 * evolve_pi src/evolve_pi.f90:1-322 (type 'RPMD', nbaro = 0)
 * ==================================================================================== */
static void outer_ostep(const rpmd_params* P, double* p, const double* mass, int na, int nb, double dt,
                        double beta, orc_rng* s, orc_gle* g) {
    if (P->therm == RPMD_THERM_PRG) orc_parinello_therm(p, mass, P->ttau, na, nb, 0.5 * dt, s, beta, 0);
    else if (P->therm == RPMD_THERM_PRL) orc_parinello_therm_loc(p, mass, P->ttau, na, nb, dt, s, beta);
    else if (P->therm == RPMD_THERM_GLE && g) orc_therm_gle(g, p, mass, na, nb, s);
}

/* the free ring-polymer update of evolve_pi, src/evolve_pi.f90:109-125: for every mode k, atom j
 * (type ic = mod(j,3): O, H, H) and component, p' = p m1 + r m2, r' = p m3 + r m4 with the
 * monodromy monod(3,4,nb) of ring_rpmd; p and r in normal-mode coordinates */
void orc_free_ring(double* p, double* r, const double* monod, int na, int nb) {
    size_t A = (size_t)3 * na;
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++) {
            int ic = j % 3;
            const double* mo = &monod[(size_t)k * 12];
            for (int c = 0; c < 3; c++) {
                size_t o = (size_t)k * A + 3 * (size_t)j + c;
                double pprime = p[o] * mo[0 * 3 + ic] + r[o] * mo[1 * 3 + ic];
                r[o] = p[o] * mo[2 * 3 + ic] + r[o] * mo[3 * 3 + ic];
                p[o] = pprime;
            }
        }
}

/* ======================================================================================
 * Barostats (src/module_barostat.f90) and their helpers: virial_ke (src/kinetic.f90:1-54),
 * pressure and pres_lj_tail (src/md_pressure.f90), center_water (src/md_tools.f90:367-390)
 * ==================================================================================== */

/* virial_ke src/kinetic.f90:1-54: centroid-virial kinetic energy, nc = 3 */
void orc_virial_ke(const double* r, const double* dvdr, const double* dvdr2, double* tv, double tvxyz[3],
                   double* tq1o, double* tq2o, double beta, int na, int nb, int nctot) {
    size_t A = (size_t)3 * na;
    double* rc = (double*)calloc(A, sizeof(double));
    for (int k = 0; k < nb; k++)
        for (size_t i = 0; i < A; i++) rc[i] = rc[i] + r[(size_t)k * A + i];
    for (size_t i = 0; i < A; i++) rc[i] = rc[i] / (double)nb;
    double tq1 = 0.0, tq2 = 0.0, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
    for (int k = 0; k < nb; k++)
        for (int j = 0; j < na; j++)
            for (int i = 0; i < 3; i++) {
                size_t o = (size_t)k * A + 3 * (size_t)j + i;
                double dr = r[o] - rc[3 * (size_t)j + i];
                double a = dr * dvdr[o], b = dr * dvdr2[o];
                tq1 = tq1 + a;
                tq2 = tq2 + b;
                t1[i] = t1[i] + a;
                t2[i] = t2[i] + b;
            }
    int nc = 3;
    tq1 = 0.5 * tq1 / (double)nb;
    tq2 = 0.5 * tq2 / (double)nb;
    for (int i = 0; i < 3; i++) {
        t1[i] = 0.5 * t1[i] / (double)nb;
        t2[i] = 0.5 * t2[i] / (double)nb;
    }
    double tc = 0.5 * (double)(3 * na - nc - nctot) / beta;
    *tv = tc + tq1 + tq2;
    for (int i = 0; i < 3; i++) tvxyz[i] = tc / 3.0 + t1[i] + t2[i];
    if (tq1o) *tq1o = tq1;
    if (tq2o) *tq2o = tq2;
    free(rc);
}

/* pres_lj_tail src/md_pressure.f90:37-56 */
double orc_pres_lj_tail(const rpmd_params* P, const double box[3], int na) {
    int nm = na / 3;
    double pi = acos(-1.0);
    double vol = box[0] * box[1] * box[2];
    double rho = (double)nm / vol;
    double prefac = (16.0 * pi * (rho * rho) * P->oo_eps * (P->oo_sig * P->oo_sig * P->oo_sig)) / 3.0;
    double x = P->oo_sig / P->rcut, x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x8 = x4 * x4, x9 = x8 * x;
    return prefac * ((2.0 / 3.0) * x9 - x3);
}

/* pressure src/md_pressure.f90:1-35 (rpmddft = 0) */
double orc_pressure(const rpmd_params* P, const double vir[9], double tv, int na, const double box[3]) {
    double w = VIR(vir, 0, 0) + VIR(vir, 1, 1) + VIR(vir, 2, 2);
    w = -w / 3.0;
    double vol = box[0] * box[1] * box[2];
    double pres = (2.0 * tv / 3.0 + w) / vol;
    double ptail = P->oo_gam == 0.0 ? orc_pres_lj_tail(P, box, na) : 0.0;
    return pres + ptail;
}

/* center_water src/md_tools.f90:367-390: per molecule the bead average of the SUM of its three
 * atoms' coordinates (not their mean) */
void orc_center_water(const double* r, double* pc, int nm, int nb) {
    double wt = 1.0 / (double)nb;
    size_t A = (size_t)9 * nm;
    for (int j = 0; j < nm; j++) {
        double c[3] = {0, 0, 0};
        for (int k = 0; k < nb; k++) {
            const double* x = r + (size_t)k * A + 9 * (size_t)j;
            for (int d = 0; d < 3; d++) c[d] = c[d] + x[d] + x[3 + d] + x[6 + d];
        }
        for (int d = 0; d < 3; d++) pc[3 * j + d] = c[d] * wt;
    }
}

/* every atom of molecule i in every bead moves by rcm(:, i) (scale(:) - 1) (src/module_barostat.f90:66-79,
 * 107-120, 176-190) */
static void shift_molecules(double* r, const double* rcm, const double sc[3], int nm, int nb) {
    size_t A = (size_t)9 * nm;
    for (int i = 0; i < nm; i++)
        for (int k = 0; k < nb; k++)
            for (int a = 0; a < 3; a++)
                for (int d = 0; d < 3; d++) {
                    double* x = r + (size_t)k * A + 9 * (size_t)i + 3 * a + d;
                    *x = *x + rcm[3 * i + d] * (sc[d] - 1.0);
                }
}

/* beren_driver, beren_baro, beren_baro_nc src/module_barostat.f90:8-122 (compi = 13531, the isothermal
 * compressibility of water in atomic units; rpmddft = 0) */
void orc_beren_driver(const rpmd_params* P, const orc_baro* B, const double vir[9], double tv, const double tvxyz[3],
                      double dt, double* r, double box[3], int na, int nb) {
    int nm = na / 3;
    double compi = 13531.0, bstat = compi / (3.0 * B->taub), sc[3];
    if (P->iamcub) {
        double pres = orc_pressure(P, vir, tv, na, box);
        double scale = fabs(1.0 - (dt * bstat) * (B->patm - pres));
        for (int d = 0; d < 3; d++) {
            box[d] = scale * box[d];
            sc[d] = scale;
        }
    } else {
        double ptail = P->oo_gam == 0.0 ? orc_pres_lj_tail(P, box, na) : 0.0;
        double vol = box[0] * box[1] * box[2], pres[3];
        for (int j = 0; j < 3; j++) {
            pres[j] = (-VIR(vir, j, j) + 2.0 * tvxyz[j]) / vol;
            pres[j] = pres[j] + ptail;
        }
        for (int d = 0; d < 3; d++) {
            sc[d] = fabs(1.0 - (dt * bstat) * (B->patm - pres[d]));
            box[d] = sc[d] * box[d];
        }
    }
    double* rcm = (double*)malloc(sizeof(double) * 3 * (nm > 0 ? nm : 1));
    orc_center_water(r, rcm, nm, nb);
    shift_molecules(r, rcm, sc, nm, nb);
    free(rcm);
}

/* full_forces as mc_baro sees it: with contraction the reference transforms its position argument
 * to normal modes and back in place (src/md_forces.f90:357,386), which mc_baro's caller keeps */
static int full_forces_inplace(const rpmd_params* P, double* r, int na, int nb, double* v, double* vir,
                               const double* z, const double box[3], double* dvdr, double* dvdr2) {
    double vew, voo, vint;
    int err = orc_full_forces(P, r, na, nb, v, &vew, &voo, &vint, vir, z, box, dvdr, dvdr2);
    if (!err && nb > 1 && (P->nbdf1 > 0 || P->nbdf2 > 0)) {
        orc_realft(r, 3 * na, nb, +1);
        orc_realft(r, 3 * na, nb, -1);
    }
    return err;
}

/* mc_baro src/module_barostat.f90:124-220: with probability 1/taub a random volume change
 * vn = vo + dv u(-1,1), molecules scaled by (vn/vo)^(1/3), accepted when
 * exp(-beta ((en - eo) + patm (vn - vo) - nm ln(vn/vo) / beta)) > u(0,1); the second definition of
 * the test (:190) is the one in effect */
int orc_mc_baro(const rpmd_params* P, orc_baro* B, double* r, double* dvdr, double* dvdr2, double* vir, double* eo,
                const double* z, double beta, double box[3], int na, int nb, orc_rng* s) {
    int nm = na / 3, err = 0;
    double thresh = 1.0 / B->taub;
    if (!(orc_ran2(s, 0.0, 1.0) < thresh)) return 0;
    size_t N = (size_t)3 * na * nb;
    err = full_forces_inplace(P, r, na, nb, eo, vir, z, box, dvdr, dvdr2);
    if (err) return err;
    B->nmove = B->nmove + 1;
    double vo = box[0] * box[1] * box[2];
    double vn = vo + B->dv * orc_ran2(s, -1.0, 1.0);
    double scale = pow(vn / vo, 1.0 / 3.0), boxn[3], sc[3] = {scale, scale, scale};
    for (int d = 0; d < 3; d++) boxn[d] = scale * box[d];
    double* rcm = (double*)malloc(sizeof(double) * 3 * (nm > 0 ? nm : 1));
    double* rn = (double*)malloc(sizeof(double) * N);
    double* dvdrn = (double*)malloc(sizeof(double) * N);
    double* dvdr2n = (double*)malloc(sizeof(double) * N);
    orc_center_water(r, rcm, nm, nb);
    memcpy(rn, r, sizeof(double) * N);
    shift_molecules(rn, rcm, sc, nm, nb);
    double en, virn[9];
    err = full_forces_inplace(P, rn, na, nb, &en, virn, z, boxn, dvdrn, dvdr2n);
    if (!err) {
        double test = (en - *eo) + B->patm * (vn - vo) - (double)nm * log(vn / vo) / beta;
        test = exp(-beta * test);
        int accept = test > orc_ran2(s, 0.0, 1.0);
        if (accept) {
            B->nacc = B->nacc + 1;
            *eo = en;
            memcpy(box, boxn, sizeof boxn);
            memcpy(r, rn, sizeof(double) * N);
            memcpy(dvdr, dvdrn, sizeof(double) * N);
            memcpy(dvdr2, dvdr2n, sizeof(double) * N);
            memcpy(vir, virn, sizeof virn);
        }
    }
    free(rcm); free(rn); free(dvdrn); free(dvdr2n);
    return err;
}

static int evolve_pi_core(const rpmd_params* P, double* p, double* r, double* v, double* vew, double* vlj,
                          double* vint, double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                          double box[3], const double* z, double beta, double* vir, double* vir_lf, orc_rng* s,
                          orc_gle* g, orc_baro* B);

int orc_evolve_pi(const rpmd_params* P, double* p, double* r, double* v, double* vew, double* vlj,
                  double* vint, double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                  const double box[3], const double* z, double beta, double* vir, double* vir_lf, orc_rng* s,
                  orc_gle* g) {
    double b[3] = {box[0], box[1], box[2]};
    return evolve_pi_core(P, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, b, z, beta, vir, vir_lf, s, g,
                          NULL);
}

/* evolve_pi with nbaro = 1: box and vir_lf in/out (vir = vir_lf + the step's high-frequency virial
 * feeds the barostat, src/evolve_pi.f90:163-189) */
int orc_evolve_pi_npt(const rpmd_params* P, double* p, double* r, double* v, double* vew, double* vlj,
                      double* vint, double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                      double box[3], const double* z, double beta, double* vir, double* vir_lf, orc_rng* s,
                      orc_gle* g, orc_baro* B) {
    return evolve_pi_core(P, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, s, g,
                          B);
}

static int evolve_pi_core(const rpmd_params* P, double* p, double* r, double* v, double* vew, double* vlj,
                          double* vint, double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                          double box[3], const double* z, double beta, double* vir, double* vir_lf, orc_rng* s,
                          orc_gle* g, orc_baro* B) {
    size_t A = (size_t)3 * na, N = A * nb;
    int mts = P->mts, nbdf1 = P->nbdf1, nbdf2 = P->nbdf2, err = 0;
    double* dvdre = (double*)calloc(N, sizeof(double));
    double* rst = (double*)calloc(N, sizeof(double));
    double* dvdrl = (double*)calloc(N, sizeof(double));
    double* monod = (double*)calloc((size_t)12 * nb, sizeof(double));
    double vir_itr[9] = {0}, vir_lj[9] = {0}, vir_ew[9] = {0}, vir_hf[9] = {0}, vir_ewc[9] = {0};
    double ve = 0.0, tv = 0.0, tvxyz[3] = {0, 0, 0};
    *vew = 0.0; *vlj = 0.0; *vint = 0.0;
    double dtsmall = dt / (double)mts, halfdtsmall = 0.5 * dtsmall, halfdt = 0.5 * dt;
    orc_ring_rpmd(mass, monod, na, nb, dtsmall, beta);
    outer_ostep(P, p, mass, na, nb, dt, beta, s, g);
    for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdt * dvdr[i];
    for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdtsmall * dvdr2[i];
    orc_realft(p, (int)A, nb, +1);
    orc_realft(r, (int)A, nb, +1);
    for (int im = 1; im <= mts && !err; im++) {
        if (P->therm == RPMD_THERM_PRO) orc_therm_pro(p, mass, beta, P->ttau, halfdtsmall, na, nb, s);
        else if (P->therm == RPMD_THERM_PRM) orc_therm_pro_global(p, mass, beta, P->ttau, halfdtsmall, na, nb, s);
        orc_free_ring(p, r, monod, na, nb);
        if (P->therm == RPMD_THERM_PRO) orc_therm_pro(p, mass, beta, P->ttau, halfdtsmall, na, nb, s);
        else if (P->therm == RPMD_THERM_PRM) orc_therm_pro_global(p, mass, beta, P->ttau, halfdtsmall, na, nb, s);
        memcpy(rst, r, sizeof(double) * N);
        orc_realft(rst, (int)A, nb, -1);
        err = orc_forces(P, rst, nb, na, box, z, 4, vint, dvdr2, vir_itr);
        if (B) {
            double tv_itr, tvxyz_itr[3];
            orc_virial_ke(rst, dvdr, dvdr2, &tv_itr, tvxyz_itr, NULL, NULL, beta, na, nb, 0);
            tv = tv + tv_itr;
            for (int d = 0; d < 3; d++) tvxyz[d] = tvxyz[d] + tvxyz_itr[d];
        }
        for (int m = 0; m < 9; m++) vir_hf[m] = vir_hf[m] + vir_itr[m];
        if (im != mts) {
            orc_realft(dvdr2, (int)A, nb, +1);
            for (size_t i = 0; i < N; i++) p[i] = p[i] - dtsmall * dvdr2[i];
        }
    }
    tv = tv / (double)mts;
    for (int d = 0; d < 3; d++) tvxyz[d] = tvxyz[d] / (double)mts;
    for (int m = 0; m < 9; m++) vir_hf[m] = vir_hf[m] / (double)mts;
    if (!err && B) {
        /* barostat on the bead positions, then the normal modes of the scaled positions
         * (src/evolve_pi.f90:163-189) */
        double vb[9];
        for (int m = 0; m < 9; m++) vb[m] = vir_lf[m] + vir_hf[m];
        if (B->baro == ORC_BARO_BER) {
            orc_beren_driver(P, B, vb, tv, tvxyz, dt, rst, box, na, nb);
        } else if (B->baro == ORC_BARO_MCI) {
            double vmc;
            err = orc_mc_baro(P, B, rst, dvdr, dvdr2, vb, &vmc, z, beta, box, na, nb, s);
        } else {
            err = -1;
        }
        memcpy(r, rst, sizeof(double) * N);
        orc_realft(r, (int)A, nb, +1);
    }
    if (!err && nbdf1 > 0) {
        err = orc_rp_contract_nm(P, r, vew, dvdr, nb, na, box, z, vir_ew, nbdf1, 2);
        if (!err && P->sig > 0.0) {
            err = orc_rp_contract_nm(P, r, &ve, dvdre, nb, na, box, z, vir_ewc, nbdf1, 5);
            *vew = *vew - ve;
            for (int m = 0; m < 9; m++) vir_ew[m] = vir_ew[m] - vir_ewc[m];
            for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] - dvdre[i];
        }
    }
    if (!err && nbdf2 > 0) err = orc_rp_contract_nm(P, r, vlj, dvdrl, nb, na, box, z, vir_lj, nbdf2, 3);
    orc_realft(p, (int)A, nb, -1);
    memcpy(r, rst, sizeof(double) * N);
    for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdtsmall * dvdr2[i];
    if (nbdf1 > 0 && nbdf2 > 0) {
        for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdrl[i];
        orc_realft(dvdr, (int)A, nb, -1);
    } else if (nbdf1 > 0) {
        orc_realft(dvdr, (int)A, nb, -1);
    } else if (nbdf2 > 0) {
        orc_realft(dvdrl, (int)A, nb, -1);
    }
    if (!err) {
        if (nbdf1 <= 0) {
            err = orc_forces(P, r, nb, na, box, z, 2, vew, dvdr, vir_ew);
        } else if (P->sig > 0.0) {
            err = orc_forces(P, r, nb, na, box, z, 5, &ve, dvdre, vir_ewc);
            *vew = *vew + ve;
            for (int m = 0; m < 9; m++) vir_ew[m] = vir_ew[m] + vir_ewc[m];
            for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdre[i];
        }
    }
    if (!err && nbdf2 <= 0) err = orc_forces(P, r, nb, na, box, z, 3, vlj, dvdrl, vir_lj);
    if (nbdf1 <= 0 || nbdf2 <= 0)
        for (size_t i = 0; i < N; i++) dvdr[i] = dvdr[i] + dvdrl[i];
    for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdt * dvdr[i];
    outer_ostep(P, p, mass, na, nb, dt, beta, s, g);
    *v = *vew + *vlj + *vint;
    for (int m = 0; m < 9; m++) {
        vir_lf[m] = vir_lj[m] + vir_ew[m];
        vir[m] = vir_lf[m] + vir_hf[m];
    }
    free(dvdre); free(rst); free(dvdrl); free(monod);
    return err;
}

/* ======================================================================================
 * Rigid molecules: RATTLE (src/md_rattle.f90) and evolve_rig_pi (src/evolve_rigid.f90)
 * ==================================================================================== */

/* rattle_setup src/md_rattle.f90:1-31: per molecule the bonds (O,H1), (O,H2), (H1,H2) with
 * squared lengths reoh^2, reoh^2, rehh^2, rehh = 2 sin(theta/2) reoh */
static void rattle_bonds(double reoh, double theta, double rcb[3], int bi[3], int bj[3]) {
    double rehh = 2.0 * sin(0.5 * theta) * reoh;
    rcb[0] = reoh * reoh; bi[0] = 0; bj[0] = 1;
    rcb[1] = reoh * reoh; bi[1] = 0; bj[1] = 2;
    rcb[2] = rehh * rehh; bi[2] = 1; bj[2] = 2;
}

/* rattle_s1 src/md_rattle.f90:33-158 on one bead r(3,na); vir(3,3) column-major is zeroed first.
 * The reference accumulates vir(1,3) as vir(1,2) - gdx dz2 (:124) and checks only the iteration
 * count of the last molecule (:144-146); both are kept. */
int orc_rattle_s1(double* r, const double* rold, double* p, int nm, const double* mass, double dt,
                  double* dvdr, double* vir, double reoh, double theta) {
    const double tol = 1e-6;
    const int itmax = 1000;
    double rcb[3];
    int bi[3], bj[3];
    rattle_bonds(reoh, theta, rcb, bi, bj);
    for (int m = 0; m < 9; m++) vir[m] = 0.0;
#define V(i, j) vir[(j - 1) * 3 + (i - 1)]
    int it = 0;
    for (int ik = 0; ik < nm; ik++) {
        int done = 0;
        it = 0;
        while (!done && it < itmax) {
            done = 1;
            for (int b = 0; b < 3; b++) {
                int ii = 3 * ik + bi[b], jj = 3 * ik + bj[b];
                double dx1 = RX(r, 0, jj) - RX(r, 0, ii);
                double dy1 = RX(r, 1, jj) - RX(r, 1, ii);
                double dz1 = RX(r, 2, jj) - RX(r, 2, ii);
                double d2 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
                double diff = rcb[b] - d2;
                if (fabs(diff) > rcb[b] * tol) {
                    double rm1 = 1.0 / mass[jj], rm2 = 1.0 / mass[ii];
                    double dx2 = RX(rold, 0, jj) - RX(rold, 0, ii);
                    double dy2 = RX(rold, 1, jj) - RX(rold, 1, ii);
                    double dz2 = RX(rold, 2, jj) - RX(rold, 2, ii);
                    double dot = dx1 * dx2 + dy1 * dy2 + dz1 * dz2;
                    double gamma = diff / (2.0 * (rm1 + rm2) * dot);
                    double gdx = gamma * dx2, gdy = gamma * dy2, gdz = gamma * dz2;
                    RX(r, 0, jj) = RX(r, 0, jj) + rm1 * gdx;
                    RX(r, 1, jj) = RX(r, 1, jj) + rm1 * gdy;
                    RX(r, 2, jj) = RX(r, 2, jj) + rm1 * gdz;
                    RX(r, 0, ii) = RX(r, 0, ii) - rm2 * gdx;
                    RX(r, 1, ii) = RX(r, 1, ii) - rm2 * gdy;
                    RX(r, 2, ii) = RX(r, 2, ii) - rm2 * gdz;
                    gdx = gdx / dt; gdy = gdy / dt; gdz = gdz / dt;
                    RX(p, 0, jj) = RX(p, 0, jj) + gdx;
                    RX(p, 1, jj) = RX(p, 1, jj) + gdy;
                    RX(p, 2, jj) = RX(p, 2, jj) + gdz;
                    RX(p, 0, ii) = RX(p, 0, ii) - gdx;
                    RX(p, 1, ii) = RX(p, 1, ii) - gdy;
                    RX(p, 2, ii) = RX(p, 2, ii) - gdz;
                    gdx = 2.0 * gdx; gdy = 2.0 * gdy; gdz = 2.0 * gdz;
                    V(1, 1) = V(1, 1) - gdx * dx2;
                    V(1, 2) = V(1, 2) - gdx * dy2;
                    V(1, 3) = V(1, 2) - gdx * dz2;
                    V(2, 3) = V(2, 3) - gdy * dz2;
                    V(2, 2) = V(2, 2) - gdy * dy2;
                    V(3, 3) = V(3, 3) - gdz * dz2;
                    RX(dvdr, 0, jj) = RX(dvdr, 0, jj) + gdx;
                    RX(dvdr, 1, jj) = RX(dvdr, 1, jj) + gdy;
                    RX(dvdr, 2, jj) = RX(dvdr, 2, jj) + gdz;
                    RX(dvdr, 0, ii) = RX(dvdr, 0, ii) - gdx;
                    RX(dvdr, 1, ii) = RX(dvdr, 1, ii) - gdy;
                    RX(dvdr, 2, ii) = RX(dvdr, 2, ii) - gdz;
                    done = 0;
                }
            }
            it = it + 1;
        }
    }
    int err = (nm > 0 && it >= itmax) ? RPMD_ERR_RATTLE : 0;
    V(2, 1) = V(1, 2);
    V(3, 1) = V(1, 3);
    V(3, 2) = V(2, 3);
    return err;
}

/* rattle_s2 src/md_rattle.f90:160-271 (velocity stage, tolerance on gamma, it <= itmax) */
int orc_rattle_s2(const double* r, double* p, int nm, const double* mass, double dt, double* dvdr,
                  double* vir, double reoh, double theta) {
    const double tol = 1e-6;
    const int itmax = 1000;
    const double halfdt = 0.5 * dt;
    double rcb[3];
    int bi[3], bj[3];
    rattle_bonds(reoh, theta, rcb, bi, bj);
    for (int m = 0; m < 9; m++) vir[m] = 0.0;
    int it = 0;
    for (int ik = 0; ik < nm; ik++) {
        int done = 0;
        it = 0;
        while (!done && it <= itmax) {
            done = 1;
            for (int b = 0; b < 3; b++) {
                int ii = 3 * ik + bi[b], jj = 3 * ik + bj[b];
                double dx1 = RX(r, 0, jj) - RX(r, 0, ii);
                double dy1 = RX(r, 1, jj) - RX(r, 1, ii);
                double dz1 = RX(r, 2, jj) - RX(r, 2, ii);
                double rm1 = 1.0 / mass[jj], rm2 = 1.0 / mass[ii];
                double dvx1 = RX(p, 0, jj) * rm1 - RX(p, 0, ii) * rm2;
                double dvy1 = RX(p, 1, jj) * rm1 - RX(p, 1, ii) * rm2;
                double dvz1 = RX(p, 2, jj) * rm1 - RX(p, 2, ii) * rm2;
                double dot = dx1 * dvx1 + dy1 * dvy1 + dz1 * dvz1;
                double gamma = -dot / ((rm1 + rm2) * rcb[b]);
                if (fabs(gamma) > tol) {
                    double gdx = gamma * dx1, gdy = gamma * dy1, gdz = gamma * dz1;
                    RX(p, 0, jj) = RX(p, 0, jj) + gdx;
                    RX(p, 1, jj) = RX(p, 1, jj) + gdy;
                    RX(p, 2, jj) = RX(p, 2, jj) + gdz;
                    RX(p, 0, ii) = RX(p, 0, ii) - gdx;
                    RX(p, 1, ii) = RX(p, 1, ii) - gdy;
                    RX(p, 2, ii) = RX(p, 2, ii) - gdz;
                    gdx = gdx / halfdt; gdy = gdy / halfdt; gdz = gdz / halfdt;
                    V(1, 1) = V(1, 1) - gdx * dx1;
                    V(1, 2) = V(1, 2) - gdx * dy1;
                    V(1, 3) = V(1, 2) - gdx * dz1;
                    V(2, 3) = V(2, 3) - gdy * dz1;
                    V(2, 2) = V(2, 2) - gdy * dy1;
                    V(3, 3) = V(3, 3) - gdz * dz1;
                    RX(dvdr, 0, jj) = RX(dvdr, 0, jj) + gdx;
                    RX(dvdr, 1, jj) = RX(dvdr, 1, jj) + gdy;
                    RX(dvdr, 2, jj) = RX(dvdr, 2, jj) + gdz;
                    RX(dvdr, 0, ii) = RX(dvdr, 0, ii) - gdx;
                    RX(dvdr, 1, ii) = RX(dvdr, 1, ii) - gdy;
                    RX(dvdr, 2, ii) = RX(dvdr, 2, ii) - gdz;
                    done = 0;
                }
            }
            it = it + 1;
        }
    }
    int err = (nm > 0 && it >= itmax) ? RPMD_ERR_RATTLE : 0;
    V(2, 1) = V(1, 2);
    V(3, 1) = V(1, 3);
    V(3, 2) = V(2, 3);
#undef V
    return err;
}

/* bead_forces src/evolve_rigid.f90:236-278: free ring-polymer spring gradient in bead space,
 * m omega_n^2 (2 x_k - x_{k-1} - x_{k+1}), omega_n = nb / beta (hbar = 1); zero for nb = 1 */
void orc_bead_forces(const double* r, int na, int nb, double* dvdrb, double beta, const double* mass) {
    size_t A = (size_t)3 * na;
    double omegan = (double)nb / (beta * 1.0);
    omegan = omegan * omegan;
    for (size_t i = 0; i < A * nb; i++) dvdrb[i] = 0.0;
    if (nb <= 1) return;
    for (int i = 0; i < na; i++)
        for (int k = 0; k < nb; k++) {
            int k2 = (k == 0) ? nb - 1 : k - 1, k3 = (k == nb - 1) ? 0 : k + 1;
            for (int c = 0; c < 3; c++) {
                double dx1 = r[k * A + 3 * i + c], dx2 = r[k2 * A + 3 * i + c], dx3 = r[k3 * A + 3 * i + c];
                dvdrb[k * A + 3 * i + c] = mass[i] * omegan * (2.0 * dx1 - dx2 - dx3);
            }
        }
}

/* evolve_rig_pi src/evolve_rigid.f90:100-234 (nbaro = 0).  The reference adds dvdr_corr, v_corr
 * and vir_corr, which it never sets there (the correction call is commented out, :178-180); they
 * are taken as zero, as evolve_rig_cl sets them (:19-21).  vir is in/out: the reference adds the
 * incoming value (:208).  nctot = 3 nm constraints enter the Bussi degrees of freedom. */
int orc_evolve_rig_pi(const rpmd_params* P, double* p, double* r, double* v, double* v_lf, double* v_hf,
                      double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb,
                      const double box[3], const double* z, double beta, double* vir, double* vir_lf,
                      orc_rng* s) {
    size_t A = (size_t)3 * na, N = A * nb;
    int nm = na / 3, nstep = nb, err = 0;
    int nctot = 3 * nm;
    double halfdt = 0.5 * dt, smalldt = dt / (double)nstep, halfsmalldt = 0.5 * smalldt;
    double* rold = (double*)malloc(sizeof(double) * A);
    double vir_hf[9] = {0}, vir_tmp[9];
    if (P->therm == RPMD_THERM_PRG) orc_parinello_therm(p, mass, P->ttau, na, nb, halfdt, s, beta, nctot);
    else if (P->therm == RPMD_THERM_PRL) orc_parinello_therm_loc(p, mass, P->ttau, na, nb, halfdt, s, beta);
    for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdt * dvdr[i];
    orc_bead_forces(r, na, nb, dvdr2, beta, mass);
    for (int is = 0; is < nstep && !err; is++) {
        for (int k = 0; k < nb && !err; k++) {
            double *rk = r + k * A, *pk = p + k * A, *dk = dvdr2 + k * A;
            for (int j = 0; j < na; j++)
                for (int i = 0; i < 3; i++) {
                    RX(rold, i, j) = RX(rk, i, j);
                    RX(pk, i, j) = RX(pk, i, j) - halfsmalldt * RX(dk, i, j);
                    RX(rk, i, j) = RX(rk, i, j) + RX(pk, i, j) * smalldt / mass[j];
                }
            err = orc_rattle_s1(rk, rold, pk, nm, mass, smalldt, dk, vir_hf, P->reoh, P->theta);
        }
        if (err) break;
        orc_bead_forces(r, na, nb, dvdr2, beta, mass);
        for (size_t i = 0; i < N; i++) p[i] = p[i] - halfsmalldt * dvdr2[i];
        for (int k = 0; k < nb && !err; k++)
            err = orc_rattle_s2(r + k * A, p + k * A, nm, mass, smalldt, dvdr2 + k * A, vir_hf, P->reoh,
                                P->theta);
    }
    if (!err) err = orc_forces(P, r, nb, na, box, z, 1, v, dvdr, vir_lf);
    if (!err) {
        for (size_t i = 0; i < N; i++) p[i] = p[i] - halfdt * dvdr[i];
        for (int m = 0; m < 9; m++) vir_hf[m] = 0.0;
        for (int k = 0; k < nb && !err; k++) {
            err = orc_rattle_s2(r + k * A, p + k * A, nm, mass, dt, dvdr + k * A, vir_tmp, P->reoh, P->theta);
            for (int m = 0; m < 9; m++) vir_hf[m] = vir_hf[m] + vir_tmp[m];
        }
    }
    if (!err) {
        if (P->therm == RPMD_THERM_PRG) orc_parinello_therm(p, mass, P->ttau, na, nb, halfdt, s, beta, nctot);
        else if (P->therm == RPMD_THERM_PRL) orc_parinello_therm_loc(p, mass, P->ttau, na, nb, halfdt, s, beta);
        for (int m = 0; m < 9; m++) {
            vir_hf[m] = vir_hf[m] / (double)nb;
            vir[m] = vir_lf[m] + vir_hf[m] + vir[m];
        }
        for (size_t i = 0; i < N; i++) dvdr2[i] = 0.0;
        *v_lf = *v;
        *v_hf = 0.0;
    }
    free(rold);
    return err;
}

/* evolve_rig_pi with nbaro = 1 (src/evolve_rigid.f90:203-220): after the rigid step (dvdr2 = 0),
 * BER from virial_ke of the final positions with the step's intermolecular gradient (nctot = 3 nm)
 * and the step's in/out virial; MCI from mc_baro, which replaces v, vir, dvdr and dvdr2 by its
 * full_forces values when it tries a move; then v_lf = v, v_hf = 0 */
int orc_evolve_rig_pi_npt(const rpmd_params* P, double* p, double* r, double* v, double* v_lf, double* v_hf,
                          double* dvdr, double* dvdr2, double dt, const double* mass, int na, int nb, double box[3],
                          const double* z, double beta, double* vir, double* vir_lf, orc_rng* s, orc_baro* B) {
    int err = orc_evolve_rig_pi(P, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, s);
    if (err) return err;
    int nm = na / 3;
    if (B->baro == ORC_BARO_BER) {
        double tv, tvxyz[3];
        orc_virial_ke(r, dvdr, dvdr2, &tv, tvxyz, NULL, NULL, beta, na, nb, 3 * nm);
        orc_beren_driver(P, B, vir, tv, tvxyz, dt, r, box, na, nb);
    } else if (B->baro == ORC_BARO_MCI) {
        err = orc_mc_baro(P, B, r, dvdr, dvdr2, vir, v, z, beta, box, na, nb, s);
    } else {
        err = -1;
    }
    *v_lf = *v;
    *v_hf = 0.0;
    return err;
}

