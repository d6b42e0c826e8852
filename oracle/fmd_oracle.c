/*
 * fmd_oracle.c -- FP64 CPU restatement of the reference FMD step path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path.  Loaded only by
 * tests/, the smoke check and bench.py's CPU-baseline leg.  Nothing in the
 * product links it.
 *
 * Parity status of each part (how it is pinned, see tests/test_oracle_*.py):
 *   spline prepare/eval    scipy CubicSpline(bc_type='natural') on the same
 *                          knots; closed-form cubic reproduction
 *   EAM r^2 tables         analytic setfl (f, phi, F closed forms) vs tables;
 *                          Cu01.eam.alloy fixture round trip
 *   LJ / Morse             two-atom closed forms; O(N^2) brute force
 *   EAM pass 1/0           O(N^2) brute force built from scipy splines;
 *                          perfect-lattice symmetry (equal rho, zero force)
 *   binning order          hand-derived orders on small cases; invariants
 *   integrators            closed-form single steps; Berendsen lambda formula
 *   timers / step counts   closed-form counts
 *   1-D TTM                hand-computed single sub-steps
 *   hybrid passes          brute force per category
 */
#include "fmd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXK 8
#define MAXE 8

/* ------------------------------------------------------------------ cells */

typedef struct
{
    int n, cap;
    double *x, *v, *F, *aux;
    int *grp;
    unsigned *aid;
    int *kind;
    int cnb0len, cnb1len, cnb2len;
    int *cnb0, *cnb1, *cnb2;     /* storage indices of neighbour cells */
    int own_cnb1, own_cnb2;
} ocell;

static void cell_reserve(ocell *c, int need)
{
    if (need <= c->cap) return;
    int cap = c->cap ? c->cap : 4;
    while (cap < need) cap *= 2;
    c->x = realloc(c->x, 3 * cap * sizeof(double));
    c->v = realloc(c->v, 3 * cap * sizeof(double));
    c->F = realloc(c->F, 3 * cap * sizeof(double));
    c->aux = realloc(c->aux, cap * sizeof(double));
    c->grp = realloc(c->grp, cap * sizeof(int));
    c->aid = realloc(c->aid, cap * sizeof(unsigned));
    c->kind = realloc(c->kind, cap * sizeof(int));
    c->cap = cap;
}

/* _fmd_cell_new_particle, cell.h:84-95 (capacity policy is irrelevant) */
static int cell_new(ocell *c)
{
    cell_reserve(c, c->n + 1);
    return c->n++;
}

/* _fmd_cell_copy_atom_from_cell_to_cell, cell.c:132-149 */
static void cell_copy(ocell *from, int i, ocell *to, int j)
{
    for (int d = 0; d < 3; d++)
    {
        to->x[3 * j + d] = from->x[3 * i + d];
        to->v[3 * j + d] = from->v[3 * i + d];
        to->F[3 * j + d] = from->F[3 * i + d];
    }
    to->aux[j] = from->aux[i];
    to->grp[j] = from->grp[i];
    to->aid[j] = from->aid[i];
    to->kind[j] = from->kind[i];
}

/* _fmd_cell_remove_atom, cell.c:151-160: swap with the last */
static void cell_remove(ocell *c, int i)
{
    c->n--;
    if (c->n != i) cell_copy(c, c->n, c, i);
}

static void cell_free(ocell *c)
{
    free(c->x); free(c->v); free(c->F); free(c->aux);
    free(c->grp); free(c->aid); free(c->kind);
    if (c->own_cnb1) free(c->cnb1);
    if (c->own_cnb2) free(c->cnb2);
    free(c->cnb0);
}

/* ------------------------------------------------------------------ state */

typedef struct
{
    int cat;
    double eps, eps4, sig, sig_sqr, cutoff_sqr;  /* LJ, lj.h:60-67 */
    double D0, alpha, r0;                         /* Morse, morse.h:60-66 */
    int iloc, jloc;                               /* EAM local element indices */
} opot;

typedef struct
{
    int loaded;
    int ne, Nrho, Nr, Nr2;
    double drho, dr, dr2, cutoff_sqr;
    char name[MAXE][32];
    double emass[MAXE], latt[MAXE];
    double *F[MAXE], *FDD[MAXE], *rho[MAXE], *rhoDD[MAXE];
    double *phi[MAXE][MAXE], *phiDD[MAXE][MAXE];
} oeam;

typedef struct
{
    int on;
    orc_ttm_cfg_t c;
    int n;                /* ncx * ncy * (ncz + 2) */
    int ncx, ncy, nzg;    /* nzg = ncz + 2 */
    int dim3;             /* ncx x ncy x ncz solver */
    double tcellh[3], tvol;
    unsigned *num;
    double *mass, *vcm, *Ti, *Te, *Te2, *xi, *Ke, *Ce, *Gc, *CDD, *GDD;
    double *T_C, *Ct, *T_G, *Gt;
    double frontsurf;
    unsigned initial_atoms, min_atoms;
} ottm;

struct orc
{
    double l[3];
    int pbc[3];
    int nk;
    double mass[MAXK];
    char names[MAXK][32];
    int elem_of_kind[MAXK];
    opot pot[MAXK][MAXK];
    oeam eam;
    int nc[3];
    double cellh[3];
    /* global grid (before distribution) */
    ocell *gg;
    int ngg;
    /* local grid with one ghost layer */
    int cn[3];            /* nc + 2 */
    int nci, ncm;         /* interior, total */
    int *gridp;           /* local 3-D index -> storage index */
    ocell *grid;
    int distributed;
    unsigned total;
    int natoms;
    double PE, FembSum, KE, T, p[3];
    int N;
    int nerr, lasterr;
    ottm ttm;
};

static int idx3(const int *n, int a, int b, int c) { return c + n[2] * (b + n[1] * a); }

static void orc_err(orc_t *o, int code) { o->nerr++; o->lasterr = code; }

orc_t *orc_create(const double l[3], const int pbc[3], int nkinds, const double *mass_amu,
                  const char *const *names)
{
    if (nkinds < 1 || nkinds > MAXK) return NULL;
    orc_t *o = calloc(1, sizeof(*o));
    for (int d = 0; d < 3; d++) { o->l[d] = l[d]; o->pbc[d] = pbc[d]; }
    o->nk = nkinds;
    for (int k = 0; k < nkinds; k++)
    {
        o->mass[k] = mass_amu[k] / ORC_MASS_UNIT;   /* potential.c:47 */
        snprintf(o->names[k], sizeof o->names[k], "%s", names ? names[k] : "X");
        o->elem_of_kind[k] = -1;
    }
    return o;
}

void orc_free(orc_t *o)
{
    if (!o) return;
    for (int i = 0; i < o->ngg; i++) cell_free(&o->gg[i]);
    free(o->gg);
    for (int i = 0; i < o->ncm; i++) cell_free(&o->grid[i]);
    free(o->grid);
    free(o->gridp);
    for (int e = 0; e < o->eam.ne; e++)
    {
        free(o->eam.F[e]); free(o->eam.FDD[e]); free(o->eam.rho[e]); free(o->eam.rhoDD[e]);
        for (int f = 0; f <= e; f++) { free(o->eam.phi[e][f]); free(o->eam.phiDD[e][f]); }
    }
    ottm *t = &o->ttm;
    free(t->num); free(t->mass); free(t->vcm); free(t->Ti); free(t->Te); free(t->Te2);
    free(t->xi); free(t->Ke); free(t->Ce); free(t->Gc); free(t->CDD); free(t->GDD);
    free(t->T_C); free(t->Ct); free(t->T_G); free(t->Gt);
    free(o);
}

/* ------------------------------------------------------------ potentials */

/* fmd_pot_lj_apply, lj.c:92-113 */
void orc_set_lj(orc_t *o, int k1, int k2, double sigma, double eps, double cutoff)
{
    opot p = {0};
    p.cat = ORC_POT_LJ;
    p.sig = sigma; p.sig_sqr = sigma * sigma; p.eps = eps; p.eps4 = 4. * eps;
    p.cutoff_sqr = cutoff * cutoff;
    o->pot[k1][k2] = o->pot[k2][k1] = p;
}

/* fmd_pot_morse_apply, morse.c:94-114 */
void orc_set_morse(orc_t *o, int k1, int k2, double D0, double alpha, double r0, double cutoff)
{
    opot p = {0};
    p.cat = ORC_POT_MORSE;
    p.D0 = D0; p.alpha = alpha; p.r0 = r0; p.cutoff_sqr = cutoff * cutoff;
    o->pot[k1][k2] = o->pot[k2][k1] = p;
}

/* SPLINE_VAL / SPLINE_DERIV, cspline.h:26-27 (same operation order) */
static inline double spl_val(double a, double b, const double *y, int lo, int hi, const double *yDD, double h)
{
    return (a * y[lo] + b * y[hi] + ((a * a * a - a) * yDD[lo] + (b * b * b - b) * yDD[hi]) * (h * h) / 6.0);
}
static inline double spl_der(double a, double b, const double *y, int lo, int hi, const double *yDD, double h)
{
    return ((y[hi] - y[lo]) / h + ((3 * b * b - 1) * yDD[hi] - (3 * a * a - 1) * yDD[lo]) * h / 6.0);
}

/* _fmd_spline_prepare, cspline.c:50-74: natural spline, uniform knots */
void orc_spline_prepare(double h, const double *y, int n, double *yDD)
{
    double *u = malloc((n > 1 ? n - 1 : 1) * sizeof(double));
    double p, sig = 0.5;
    yDD[0] = u[0] = 0.0;
    for (int i = 1; i <= n - 2; i++)
    {
        p = sig * yDD[i - 1] + 2.0;
        yDD[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - 2 * y[i] + y[i - 1]) / h;
        u[i] = (3.0 * u[i] / h - sig * u[i - 1]) / p;
    }
    yDD[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; k--) yDD[k] = yDD[k] * yDD[k + 1] + u[k];
    free(u);
}

/* _fmd_spline_prepare_adv, cspline.c:24-48: natural spline, general knots */
void orc_spline_prepare_adv(const double *x, const double *y, int n, double *yDD)
{
    double *u = malloc((n > 1 ? n - 1 : 1) * sizeof(double));
    double p, sig;
    yDD[0] = u[0] = 0.0;
    for (int i = 1; i <= n - 2; i++)
    {
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        p = sig * yDD[i - 1] + 2.0;
        yDD[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    yDD[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; k--) yDD[k] = yDD[k] * yDD[k + 1] + u[k];
    free(u);
}

/* _fmd_spline_val / _fmd_spline_deriv, cspline.c:76-103 */
double orc_spline_val(double h, const double *ya, const double *yDDa, double x)
{
    int lo = (int)(x / h), hi = lo + 1;
    double a = hi - x / h, b = 1 - a;
    return spl_val(a, b, ya, lo, hi, yDDa, h);
}
double orc_spline_deriv(double h, const double *ya, const double *yDDa, double x)
{
    int lo = (int)(x / h), hi = lo + 1;
    double a = hi - x / h, b = 1 - a;
    return spl_der(a, b, ya, lo, hi, yDDa, h);
}

/* EAM_convert_r_to_r2, eam.c:324-341.  At the last point sqrt(i*dr2)/dr can
   reach Nr-1 exactly; _fmd_spline_val (cspline.c:76-89) then reads element
   Nr, one past the table, with weight b = 0.  A zero pad element keeps that
   read inside the buffer and leaves the value unchanged. */
static void eam_r_to_r2(oeam *e, const double *src, double *dst)
{
    double *y = calloc(e->Nr + 1, sizeof(double));
    double *dd = calloc(e->Nr + 1, sizeof(double));
    memcpy(y, src, e->Nr * sizeof(double));
    orc_spline_prepare(e->dr, y, e->Nr, dd);
    for (int i = 0; i < e->Nr2; i++) dst[i] = orc_spline_val(e->dr, y, dd, sqrt(i * e->dr2));
    free(y);
    free(dd);
}

/* load_DYNAMOsetfl, eam.c:376-555 (single rank: no broadcast) */
int orc_load_eam(orc_t *o, const char *path)
{
    oeam *e = &o->eam;
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char str[1024];
    for (int i = 0; i < 3; i++)
        if (!fgets(str, sizeof str, fp)) { fclose(fp); return -2; }
    if (fscanf(fp, "%d", &e->ne) != 1 || e->ne < 1 || e->ne > MAXE) { fclose(fp); return -2; }
    for (int i = 0; i < e->ne; i++)
        if (fscanf(fp, "%31s", e->name[i]) != 1) { fclose(fp); return -2; }
    double cutoff;
    if (fscanf(fp, "%d%lf%d%lf%lf", &e->Nrho, &e->drho, &e->Nr, &e->dr, &cutoff) != 5) { fclose(fp); return -2; }
    e->Nr2 = (e->Nr += 2);
    if ((e->Nr - 1) * e->dr <= cutoff) { fclose(fp); return -2; }
    e->cutoff_sqr = cutoff * cutoff;
    e->dr2 = ((e->Nr - 1) * e->dr) * ((e->Nr - 1) * e->dr) / (e->Nr2 - 1);
    double *tmp = malloc(e->Nr * sizeof(double));
    for (int i = 0; i < e->ne; i++)
    {
        if (fscanf(fp, "%1023s%lf%lf%1023s", str, &e->emass[i], &e->latt[i], str) != 4) { fclose(fp); free(tmp); return -2; }
        e->emass[i] /= ORC_MASS_UNIT;
        e->F[i] = malloc(e->Nrho * sizeof(double));
        for (int j = 0; j < e->Nrho; j++)
            if (fscanf(fp, "%lf", &e->F[i][j]) != 1) { fclose(fp); free(tmp); return -2; }
        e->rho[i] = malloc(e->Nr2 * sizeof(double));
        for (int j = 0; j < e->Nr - 2; j++)
            if (fscanf(fp, "%lf", &tmp[j]) != 1) { fclose(fp); free(tmp); return -2; }
        tmp[e->Nr - 1] = tmp[e->Nr - 2] = 0.;
        eam_r_to_r2(e, tmp, e->rho[i]);
        e->FDD[i] = malloc(e->Nrho * sizeof(double));
        e->rhoDD[i] = malloc(e->Nr2 * sizeof(double));
        orc_spline_prepare(e->drho, e->F[i], e->Nrho, e->FDD[i]);
        orc_spline_prepare(e->dr2, e->rho[i], e->Nr2, e->rhoDD[i]);
    }
    for (int i = 0; i < e->ne; i++)
        for (int j = 0; j <= i; j++)
        {
            for (int k = 0; k < e->Nr - 2; k++)
            {
                if (fscanf(fp, "%lf", &tmp[k]) != 1) { fclose(fp); free(tmp); return -2; }
                if (k == 0) tmp[k] = FLT_MAX;
                else tmp[k] /= k * e->dr;
            }
            tmp[e->Nr - 1] = tmp[e->Nr - 2] = 0.;
            e->phi[i][j] = malloc(e->Nr2 * sizeof(double));
            eam_r_to_r2(e, tmp, e->phi[i][j]);
            e->phi[j][i] = e->phi[i][j];
            e->phiDD[i][j] = malloc(e->Nr2 * sizeof(double));
            orc_spline_prepare(e->dr2, e->phi[i][j], e->Nr2, e->phiDD[i][j]);
            e->phiDD[j][i] = e->phiDD[i][j];
        }
    free(tmp);
    fclose(fp);
    e->loaded = 1;
    return 0;
}

double orc_eam_cutoff(orc_t *o) { return sqrt(o->eam.cutoff_sqr); }

/* fmd_pot_apply for EAM, potential.c:135-175 */
int orc_apply_eam(orc_t *o, int k1, int k2)
{
    oeam *e = &o->eam;
    int l1 = -1, l2 = -1;
    for (int i = 0; i < e->ne; i++)
    {
        if (strcmp(o->names[k1], e->name[i]) == 0 && l1 < 0) l1 = i;
        if (strcmp(o->names[k2], e->name[i]) == 0 && l2 < 0) l2 = i;
    }
    if (l1 < 0 || l2 < 0) return -1;
    opot p = {0};
    p.cat = ORC_POT_EAM;
    p.cutoff_sqr = e->cutoff_sqr;
    p.iloc = l1; p.jloc = l2;
    o->pot[k1][k2] = p;
    p.iloc = l2; p.jloc = l1;
    o->pot[k2][k1] = p;
    o->elem_of_kind[k1] = l1;
    o->elem_of_kind[k2] = l2;
    return 0;
}

int orc_eam_sizes(orc_t *o, int *nelem, int *nrho, int *nr2, double *drho, double *dr2, double *rc2)
{
    oeam *e = &o->eam;
    if (!e->loaded) return -1;
    *nelem = e->ne; *nrho = e->Nrho; *nr2 = e->Nr2; *drho = e->drho; *dr2 = e->dr2; *rc2 = e->cutoff_sqr;
    return 0;
}

void orc_eam_tables(orc_t *o, int el, double *F, double *FDD, double *rho, double *rhoDD)
{
    oeam *e = &o->eam;
    memcpy(F, e->F[el], e->Nrho * sizeof(double));
    memcpy(FDD, e->FDD[el], e->Nrho * sizeof(double));
    memcpy(rho, e->rho[el], e->Nr2 * sizeof(double));
    memcpy(rhoDD, e->rhoDD[el], e->Nr2 * sizeof(double));
}

void orc_eam_phi(orc_t *o, int e1, int e2, double *phi, double *phiDD)
{
    oeam *e = &o->eam;
    memcpy(phi, e->phi[e1][e2], e->Nr2 * sizeof(double));
    memcpy(phiDD, e->phiDD[e1][e2], e->Nr2 * sizeof(double));
}

/* _fmd_pot_get_largest_cutoff, potential.c:251-291 */
static double largest_cutoff(orc_t *o)
{
    double m = 0.0;
    for (int i = 0; i < o->nk; i++)
        for (int j = 0; j <= i; j++)
            if (o->pot[i][j].cutoff_sqr > m) m = o->pot[i][j].cutoff_sqr;
    return sqrt(m);
}

/* ------------------------------------------------------------------ grid */

/* _fmd_createGlobalGrid, misc.c:41-62; _fmd_subd_init for ns=1, subdomain.c:224-305 */
int orc_make_grid(orc_t *o)
{
    double rc = largest_cutoff(o);
    if (rc <= 0) return -1;
    for (int d = 0; d < 3; d++)
    {
        o->nc[d] = (int)(o->l[d] / rc);
        if (o->nc[d] < 1) return -1;
        o->cellh[d] = o->l[d] / o->nc[d];
        o->cn[d] = o->nc[d] + 2;
    }
    o->ngg = o->nc[0] * o->nc[1] * o->nc[2];
    o->gg = calloc(o->ngg, sizeof(ocell));
    o->nci = o->ngg;
    o->ncm = o->cn[0] * o->cn[1] * o->cn[2];
    o->grid = calloc(o->ncm, sizeof(ocell));
    o->gridp = malloc(o->ncm * sizeof(int));

    /* init_gridp, subdomain.c:201-222: interior first, then margin, LOOP3D order */
    int imarg = o->nci, inon = 0;
    int ic[3];
    for (ic[0] = 0; ic[0] < o->cn[0]; ic[0]++)
        for (ic[1] = 0; ic[1] < o->cn[1]; ic[1]++)
            for (ic[2] = 0; ic[2] < o->cn[2]; ic[2]++)
            {
                int margin = 0;
                for (int d = 0; d < 3; d++)
                    if (ic[d] < 1 || ic[d] >= o->nc[d] + 1) margin = 1;
                o->gridp[idx3(o->cn, ic[0], ic[1], ic[2])] = margin ? imarg++ : inon++;
            }

    /* init_cneighbs, subdomain.c:76-199 */
    for (ic[0] = 0; ic[0] < o->cn[0]; ic[0]++)
        for (ic[1] = 0; ic[1] < o->cn[1]; ic[1]++)
            for (ic[2] = 0; ic[2] < o->cn[2]; ic[2]++)
            {
                int margin = 0;
                for (int d = 0; d < 3; d++)
                    if (ic[d] < 1 || ic[d] >= o->nc[d] + 1) margin = 1;
                ocell *c = &o->grid[o->gridp[idx3(o->cn, ic[0], ic[1], ic[2])]];
                if (margin)
                {
                    c->cnb0 = malloc(27 * sizeof(int));
                    for (int j = 0; j < 27; j++)
                    {
                        int jc[3] = {j / 9 + ic[0] - 1, (j / 3) % 3 + ic[1] - 1, j % 3 + ic[2] - 1};
                        int inside = 1;
                        for (int d = 0; d < 3; d++)
                            if (jc[d] < 1 || jc[d] >= o->nc[d] + 1) inside = 0;
                        if (inside)
                        {
                            c->cnb0[c->cnb0len++] = o->gridp[idx3(o->cn, jc[0], jc[1], jc[2])];
                            if (j < 13) c->cnb2len++;
                            else c->cnb1len++;
                        }
                    }
                    c->cnb1 = c->cnb0 + c->cnb2len;
                    c->cnb2 = c->cnb0;
                }
                else
                {
                    c->cnb1len = 13;
                    c->cnb1 = malloc(13 * sizeof(int));
                    c->own_cnb1 = 1;
                    c->cnb2 = malloc(13 * sizeof(int));
                    c->own_cnb2 = 1;
                    int i = 0;
                    for (int j = 14; j < 27; j++)
                    {
                        int jc[3] = {j / 9 + ic[0] - 1, (j / 3) % 3 + ic[1] - 1, j % 3 + ic[2] - 1};
                        int inside = 1;
                        for (int d = 0; d < 3; d++)
                            if (jc[d] < 1 || jc[d] >= o->nc[d] + 1) inside = 0;
                        int s = o->gridp[idx3(o->cn, jc[0], jc[1], jc[2])];
                        c->cnb1[i++] = s;
                        if (inside) c->cnb2[c->cnb2len++] = s;
                    }
                }
            }
    return 0;
}

void orc_grid_info(orc_t *o, int nc[3], double cellh[3])
{
    for (int d = 0; d < 3; d++) { nc[d] = o->nc[d]; cellh[d] = o->cellh[d]; }
}

/* fmd_io_loadState binning, misc.c:242-259 */
int orc_add_atoms(orc_t *o, int n, const double *x, const double *v, const int *group, const int *kind)
{
    for (int i = 0; i < n; i++)
    {
        int ic[3];
        for (int d = 0; d < 3; d++)
        {
            ic[d] = (int)floor(x[3 * i + d] / o->cellh[d]);
            if (ic[d] < 0 || ic[d] >= o->nc[d]) return -1;
        }
        ocell *c = &o->gg[idx3(o->nc, ic[0], ic[1], ic[2])];
        int j = cell_new(c);
        for (int d = 0; d < 3; d++)
        {
            c->x[3 * j + d] = x[3 * i + d];
            c->v[3 * j + d] = v ? v[3 * i + d] : 0.0;
            c->F[3 * j + d] = 0.0;
        }
        c->aux[j] = 0.0;
        c->grp[j] = group ? group[i] : 0;
        c->kind[j] = kind ? kind[i] : 0;
        c->aid[j] = o->total++;
    }
    return 0;
}

/* transfer_from_globalgrid_to_rank0_grid, matter.c:329-363 */
void orc_distribute(orc_t *o)
{
    int ic[3];
    o->natoms = 0;
    for (ic[0] = 1; ic[0] < o->nc[0] + 1; ic[0]++)
        for (ic[1] = 1; ic[1] < o->nc[1] + 1; ic[1]++)
            for (ic[2] = 1; ic[2] < o->nc[2] + 1; ic[2]++)
            {
                ocell *cl = &o->grid[o->gridp[idx3(o->cn, ic[0], ic[1], ic[2])]];
                ocell *cg = &o->gg[idx3(o->nc, ic[0] - 1, ic[1] - 1, ic[2] - 1)];
                cell_reserve(cl, cg->n);
                cl->n = cg->n;
                for (int i = 0; i < cg->n; i++) cell_copy(cg, i, cl, i);
                o->natoms += cg->n;
            }
    o->distributed = 1;
}

int orc_natoms(orc_t *o)
{
    int n = 0;
    for (int i = 0; i < o->nci; i++) n += o->grid[i].n;
    return n;
}

void orc_cell_counts(orc_t *o, int *counts)
{
    for (int i = 0; i < o->nci; i++) counts[i] = o->grid[i].n;
}

void orc_get_atoms(orc_t *o, double *x, double *v, double *f, double *aux, int *group,
                   unsigned *atomid, int *kind, int *cellidx)
{
    int k = 0;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++, k++)
        {
            for (int d = 0; d < 3; d++)
            {
                if (x) x[3 * k + d] = c->x[3 * i + d];
                if (v) v[3 * k + d] = c->v[3 * i + d];
                if (f) f[3 * k + d] = c->F[3 * i + d];
            }
            if (aux) aux[k] = c->aux[i];
            if (group) group[k] = c->grp[i];
            if (atomid) atomid[k] = c->aid[i];
            if (kind) kind[k] = c->kind[i];
            if (cellidx) cellidx[k] = ic;
        }
    }
}

/* ------------------------------------------------------------ ghost layer */

typedef void (*ccopy_fn)(orc_t *o, ocell *src, ocell *dst, int dim, int dir);

/* copy_local, md-ghost.c:146-167 */
static void copy_local(orc_t *o, const int *lo, const int *hi, int dim, int dir, ccopy_fn f)
{
    int s[3];
    for (s[0] = lo[0]; s[0] < hi[0]; s[0]++)
        for (s[1] = lo[1]; s[1] < hi[1]; s[1]++)
            for (s[2] = lo[2]; s[2] < hi[2]; s[2]++)
            {
                int t[3] = {s[0], s[1], s[2]};
                t[dim] = dir < 0 ? s[dim] + o->nc[dim] : s[dim] - o->nc[dim];
                f(o, &o->grid[o->gridp[idx3(o->cn, s[0], s[1], s[2])]],
                  &o->grid[o->gridp[idx3(o->cn, t[0], t[1], t[2])]], dim, dir);
            }
}

/* ccopy_for_migrate, md-ghost.c:109-143 */
static void cc_migrate(orc_t *o, ocell *src, ocell *dst, int dim, int dir)
{
    if (src->n == 0) return;
    int old = dst->n;
    cell_reserve(dst, old + src->n);
    for (int i = 0; i < src->n; i++)
    {
        cell_copy(src, i, dst, old + i);
        dst->x[3 * (old + i) + dim] += dir > 0 ? -o->l[dim] : o->l[dim];
    }
    dst->n = old + src->n;
    src->n = 0;
}

/* ccopy_for_ghostinit, md-ghost.c:85-107 */
static void cc_ghostinit(orc_t *o, ocell *src, ocell *dst, int dim, int dir)
{
    cell_reserve(dst, src->n);
    dst->n = src->n;
    double value = dir > 0 ? -o->l[dim] : o->l[dim];
    for (int i = 0; i < src->n; i++)
    {
        for (int d = 0; d < 3; d++) dst->x[3 * i + d] = src->x[3 * i + d];
        dst->grp[i] = src->grp[i];
        dst->kind[i] = src->kind[i];
        dst->x[3 * i + dim] += value;
    }
}

/* ccopy_for_Femb, md-ghost.c:78-83 */
static void cc_femb(orc_t *o, ocell *src, ocell *dst, int dim, int dir)
{
    (void)o; (void)dim; (void)dir;
    for (int i = 0; i < src->n; i++) dst->aux[i] = src->aux[i];
}

/* ccopy_for_partialforces, md-ghost.c:69-76 */
static void cc_pforces(orc_t *o, ocell *src, ocell *dst, int dim, int dir)
{
    (void)o; (void)dim; (void)dir;
    for (int i = 0; i < src->n; i++)
        for (int d = 0; d < 3; d++) dst->F[3 * i + d] += src->F[3 * i + d];
}

/* index ranges: set_cell_indexes_for_transfer_in_direction_d, md-ghost.c:664-709.
   which=0: interior face slabs (send side of ghost init / Femb),
   which=1: ghost slabs (send side of partial-force fold-back) */
static void transfer_ranges(orc_t *o, int d, int which, int lo_l[3], int hi_l[3], int lo_u[3], int hi_u[3])
{
    for (int dd = 0; dd < 3; dd++)
    {
        if (dd == d)
        {
            if (which == 0) { lo_l[dd] = 1; hi_l[dd] = 2; lo_u[dd] = o->nc[dd]; hi_u[dd] = o->nc[dd] + 1; }
            else { lo_l[dd] = 0; hi_l[dd] = 1; lo_u[dd] = o->nc[dd] + 1; hi_u[dd] = o->nc[dd] + 2; }
        }
        else if (dd > d) { lo_l[dd] = lo_u[dd] = 0; hi_l[dd] = hi_u[dd] = o->nc[dd] + 2; }
        else { lo_l[dd] = lo_u[dd] = 1; hi_l[dd] = hi_u[dd] = o->nc[dd] + 1; }
    }
}

/* set_cell_indexes_for_migration_in_direction_d, md-ghost.c:623-662 (send = ghost slabs) */
static void migrate_ranges(orc_t *o, int d, int lo_l[3], int hi_l[3], int lo_u[3], int hi_u[3])
{
    for (int dd = 0; dd < 3; dd++)
    {
        if (dd == d) { lo_l[dd] = 0; hi_l[dd] = 1; lo_u[dd] = o->nc[dd] + 1; hi_u[dd] = o->nc[dd] + 2; }
        else if (dd > d) { lo_l[dd] = lo_u[dd] = 0; hi_l[dd] = hi_u[dd] = o->nc[dd] + 2; }
        else { lo_l[dd] = lo_u[dd] = 1; hi_l[dd] = hi_u[dd] = o->nc[dd] + 1; }
    }
}

/* cleanGridSegment, md-ghost.c:41-61 */
static int clean_segment(orc_t *o, const int *lo, const int *hi)
{
    int s[3], cnt = 0;
    for (s[0] = lo[0]; s[0] < hi[0]; s[0]++)
        for (s[1] = lo[1]; s[1] < hi[1]; s[1]++)
            for (s[2] = lo[2]; s[2] < hi[2]; s[2]++)
            {
                ocell *c = &o->grid[o->gridp[idx3(o->cn, s[0], s[1], s[2])]];
                cnt += c->n;
                c->n = 0;
            }
    return cnt;
}

/* _fmd_particles_migrate, md-ghost.c:827-869 (ns=1 branches) */
static void migrate(orc_t *o)
{
    for (int d = 0; d < 3; d++)
    {
        int lo_l[3], hi_l[3], lo_u[3], hi_u[3];
        migrate_ranges(o, d, lo_l, hi_l, lo_u, hi_u);
        if (o->pbc[d])
        {
            copy_local(o, lo_l, hi_l, d, -1, cc_migrate);
            copy_local(o, lo_u, hi_u, d, +1, cc_migrate);
        }
        else
        {
            o->natoms -= clean_segment(o, lo_l, hi_l);
            o->natoms -= clean_segment(o, lo_u, hi_u);
        }
    }
}

/* _fmd_refreshGrid, misc.c:268-308 */
int orc_refresh(orc_t *o)
{
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        int i = 0;
        while (i < c->n)
        {
            int j[3];
            for (int d = 0; d < 3; d++)
            {
                j[d] = (int)floor(c->x[3 * i + d] / o->cellh[d]) + 1;
                if (j[d] < 0 || j[d] >= o->cn[d])
                {
                    orc_err(o, 1);
                    return -1;   /* the reference returns from refresh here, misc.c:286-290 */
                }
            }
            ocell *c2 = &o->grid[o->gridp[idx3(o->cn, j[0], j[1], j[2])]];
            if (c2 != c)
            {
                int k = cell_new(c2);
                cell_copy(c, i, c2, k);
                cell_remove(c, i);
            }
            else
                i++;
        }
    }
    migrate(o);
    return 0;
}

/* _fmd_ghostparticles_init, md-ghost.c:750-786 */
static void ghost_init(orc_t *o)
{
    for (int d = 2; d >= 0; d--)
    {
        int lo_l[3], hi_l[3], lo_u[3], hi_u[3];
        transfer_ranges(o, d, 0, lo_l, hi_l, lo_u, hi_u);
        if (o->pbc[d])
        {
            copy_local(o, lo_l, hi_l, d, -1, cc_ghostinit);
            copy_local(o, lo_u, hi_u, d, +1, cc_ghostinit);
        }
    }
}

/* _fmd_ghostparticles_transfer_Femb, md-ghost.c:711-748 */
static void ghost_femb(orc_t *o)
{
    for (int d = 2; d >= 0; d--)
    {
        int lo_l[3], hi_l[3], lo_u[3], hi_u[3];
        transfer_ranges(o, d, 0, lo_l, hi_l, lo_u, hi_u);
        if (o->pbc[d])
        {
            copy_local(o, lo_l, hi_l, d, -1, cc_femb);
            copy_local(o, lo_u, hi_u, d, +1, cc_femb);
        }
    }
}

/* _fmd_ghostparticles_transfer_partialforces, md-ghost.c:788-825 */
static void ghost_fold(orc_t *o)
{
    for (int d = 0; d < 3; d++)
    {
        int lo_l[3], hi_l[3], lo_u[3], hi_u[3];
        transfer_ranges(o, d, 1, lo_l, hi_l, lo_u, hi_u);
        if (o->pbc[d])
        {
            copy_local(o, lo_l, hi_l, d, -1, cc_pforces);
            copy_local(o, lo_u, hi_u, d, +1, cc_pforces);
        }
    }
}

/* ---------------------------------------------------------------- forces */

#define ACTIVE(ag, c, i) ((ag) == -1 || (c)->grp[i] == (ag))

static inline double rv_r2(const double *x1, const double *x2, double rv[3])
{
    /* COMPUTE_rv_AND_r2, forces.h:25-34 */
    double r2 = 0.0;
    for (int d = 0; d < 3; d++)
    {
        rv[d] = x1[d] - x2[d];
        r2 += rv[d] * rv[d];
    }
    return r2;
}

/* LJ_PAIR_UPDATE_FORCE_AND_POTENERGY, lj.h:26-58 */
static inline void pair_lj(const opot *p, ocell *c1, int i1, ocell *c2, int i2, double *pe)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < p->cutoff_sqr)
    {
        double inv_r2 = 1.0 / r2;
        double inv_rs2 = p->sig_sqr * inv_r2;
        double inv_rs6 = inv_rs2 * inv_rs2 * inv_rs2;
        double inv_rs12 = inv_rs6 * inv_rs6;
        double factor = 12.0 * p->eps4 * inv_r2 * (inv_rs12 - 0.5 * inv_rs6);
        for (int d = 0; d < 3; d++)
        {
            double tmp = rv[d] * factor;
            c1->F[3 * i1 + d] += tmp;
            c2->F[3 * i2 + d] -= tmp;
        }
        *pe += p->eps4 * (inv_rs12 - inv_rs6);
    }
}

/* MORSE_PAIR_UPDATE_FORCE_AND_POTENERGY, morse.h:26-58 */
static inline void pair_morse(const opot *p, ocell *c1, int i1, ocell *c2, int i2, double *pe)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < p->cutoff_sqr)
    {
        double r = sqrt(r2);
        double inv_r = 1.0 / r;
        double exp1 = exp(-p->alpha * (r - p->r0));
        double exp2 = exp1 * exp1;
        double factor = 2.0 * p->alpha * p->D0 * inv_r * (exp2 - exp1);
        for (int d = 0; d < 3; d++)
        {
            double tmp = factor * rv[d];
            c1->F[3 * i1 + d] += tmp;
            c2->F[3 * i2 + d] -= tmp;
        }
        *pe += p->D0 * (exp2 - 2.0 * exp1);
    }
}

#define EAM_SETUP(e, r2)                       \
    double h = (e)->dr2;                       \
    int ir2 = (int)((r2) / h);                 \
    int ir2_h = ir2 + 1;                       \
    double a = ir2_h - (r2) / h;               \
    double b = 1 - a

/* EAM_PAIR_UPDATE_FORCE, eam.h:29-62 */
static inline void eam_force(oeam *e, const opot *p, ocell *c1, int i1, ocell *c2, int i2)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < e->cutoff_sqr)
    {
        EAM_SETUP(e, r2);
        double rho_ip = spl_der(a, b, e->rho[p->iloc], ir2, ir2_h, e->rhoDD[p->iloc], h);
        double mag = 2 * c2->aux[i2] * rho_ip;
        for (int d = 0; d < 3; d++)
        {
            double tmp = mag * rv[d];
            c1->F[3 * i1 + d] -= tmp;
            c2->F[3 * i2 + d] += tmp;
        }
    }
}

/* EAM_PAIR_UPDATE_FORCE_AND_POTENERGY1, eam.h:97-137 */
static inline void eam_force_pe1(oeam *e, const opot *p, ocell *c1, int i1, ocell *c2, int i2, double *pe)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < e->cutoff_sqr)
    {
        EAM_SETUP(e, r2);
        const double *phi = e->phi[p->iloc][p->jloc], *phiDD = e->phiDD[p->iloc][p->jloc];
        double phi_deriv = spl_der(a, b, phi, ir2, ir2_h, phiDD, h);
        double rho_ip = spl_der(a, b, e->rho[p->iloc], ir2, ir2_h, e->rhoDD[p->iloc], h);
        double mag = 2 * (c2->aux[i2] * rho_ip + phi_deriv);
        *pe += spl_val(a, b, phi, ir2, ir2_h, phiDD, h);
        for (int d = 0; d < 3; d++)
        {
            double tmp = mag * rv[d];
            c1->F[3 * i1 + d] -= tmp;
            c2->F[3 * i2 + d] += tmp;
        }
    }
}

/* EAM_PAIR_UPDATE_FORCE_AND_POTENERGY2, eam.h:176-228 */
static inline void eam_force_pe2(oeam *e, const opot *p, ocell *c1, int i1, ocell *c2, int i2, double *pe)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < e->cutoff_sqr)
    {
        EAM_SETUP(e, r2);
        const double *phi = e->phi[p->iloc][p->jloc], *phiDD = e->phiDD[p->iloc][p->jloc];
        double phi_deriv = spl_der(a, b, phi, ir2, ir2_h, phiDD, h);
        double rho_ip = spl_der(a, b, e->rho[p->iloc], ir2, ir2_h, e->rhoDD[p->iloc], h);
        double rho_jp;
        if (p->jloc == p->iloc) rho_jp = rho_ip;
        else rho_jp = spl_der(a, b, e->rho[p->jloc], ir2, ir2_h, e->rhoDD[p->jloc], h);
        double mag = 2 * (c1->aux[i1] * rho_jp + c2->aux[i2] * rho_ip + phi_deriv);
        *pe += spl_val(a, b, phi, ir2, ir2_h, phiDD, h);
        for (int d = 0; d < 3; d++)
        {
            double tmp = mag * rv[d];
            c1->F[3 * i1 + d] -= tmp;
            c2->F[3 * i2 + d] += tmp;
        }
    }
}

/* EAM_PAIR_UPDATE_rho_host2, eam.h:270-310 */
static inline void eam_rho2(oeam *e, const opot *p, ocell *c1, int i1, ocell *c2, int i2)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < e->cutoff_sqr)
    {
        EAM_SETUP(e, r2);
        if (p->iloc == p->jloc)
        {
            double tmp = spl_val(a, b, e->rho[p->jloc], ir2, ir2_h, e->rhoDD[p->jloc], h);
            c1->aux[i1] += tmp;
            c2->aux[i2] += tmp;
        }
        else
        {
            c1->aux[i1] += spl_val(a, b, e->rho[p->jloc], ir2, ir2_h, e->rhoDD[p->jloc], h);
            c2->aux[i2] += spl_val(a, b, e->rho[p->iloc], ir2, ir2_h, e->rhoDD[p->iloc], h);
        }
    }
}

/* EAM_PAIR_UPDATE_rho_host1, eam.h:354-379 */
static inline void eam_rho1(oeam *e, const opot *p, ocell *c1, int i1, ocell *c2, int i2)
{
    double rv[3];
    double r2 = rv_r2(&c1->x[3 * i1], &c2->x[3 * i2], rv);
    if (r2 < e->cutoff_sqr)
    {
        EAM_SETUP(e, r2);
        c2->aux[i2] += spl_val(a, b, e->rho[p->iloc], ir2, ir2_h, e->rhoDD[p->iloc], h);
    }
}

/* _fmd_clean_forces, forces.c:384-393 */
static void clean_forces(orc_t *o)
{
    for (int ic = 0; ic < o->ncm; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < 3 * c->n; i++) c->F[i] = 0.;
    }
}

/* _fmd_clean_vaream, eam.c:148-157 */
static void clean_aux(orc_t *o)
{
    for (int ic = 0; ic < o->nci; ic++)
        for (int i = 0; i < o->grid[ic].n; i++) o->grid[ic].aux[i] = 0.;
}

/* _fmd_calcFembPrime, eam.c:160-223 */
static double calc_femb_prime(orc_t *o, int ag)
{
    oeam *e = &o->eam;
    double sum = 0.;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            int el = o->elem_of_kind[c->kind[i]];
            if (el < 0) continue;
            double h = e->drho;
            double rho_host = c->aux[i];
            int irho = (int)(rho_host / h);
            if (irho >= e->Nrho - 1 || irho < 0)
            {
                orc_err(o, 5);
                irho = irho < 0 ? 0 : e->Nrho - 2;   /* the reference reads out of bounds here */
            }
            int irho_h = irho + 1;
            double a = irho_h - rho_host / h;
            double b = 1 - a;
            c->aux[i] = spl_der(a, b, e->F[el], irho, irho_h, e->FDD[el], h);
            sum += spl_val(a, b, e->F[el], irho, irho_h, e->FDD[el], h);
        }
    }
    return sum;
}

static int potcats(orc_t *o, int *has)
{
    int n = 0;
    has[0] = has[1] = has[2] = has[3] = 0;
    for (int i = 0; i < o->nk; i++)
        for (int j = 0; j <= i; j++)
            if (!has[o->pot[i][j].cat]) { has[o->pot[i][j].cat] = 1; n++; }
    return n;
}

/* _fmd_computeLJ / _fmd_computeMorse, lj.c:28-90, morse.c:30-92 */
static void compute_pairwise(orc_t *o, int ag, int cat)
{
    double pe = 0.0;
    clean_forces(o);
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c1 = &o->grid[ic];
        for (int i1 = 0; i1 < c1->n; i1++)
        {
            if (!ACTIVE(ag, c1, i1)) continue;
            int k1 = c1->kind[i1];
            for (int jc = 0; jc < c1->cnb1len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb1[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    const opot *p = &o->pot[k1][c2->kind[i2]];
                    if (cat == ORC_POT_LJ) pair_lj(p, c1, i1, c2, i2, &pe);
                    else pair_morse(p, c1, i1, c2, i2, &pe);
                }
            }
            for (int i2 = 0; i2 < i1; i2++)
            {
                if (!ACTIVE(ag, c1, i2)) continue;
                const opot *p = &o->pot[k1][c1->kind[i2]];
                if (cat == ORC_POT_LJ) pair_lj(p, c1, i1, c1, i2, &pe);
                else pair_morse(p, c1, i1, c1, i2, &pe);
            }
        }
    }
    o->PE = pe;
    o->FembSum = 0.0;
}

/* _fmd_computeEAM_pass1 (hybrid=0) and compute_hybrid_pass1 (hybrid=1),
   eam.c:225-322, forces.c:33-139 */
static void compute_pass1(orc_t *o, int ag, int hybrid)
{
    oeam *e = &o->eam;
    clean_aux(o);
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c1 = &o->grid[ic];
        for (int i1 = 0; i1 < c1->n; i1++)
        {
            if (!ACTIVE(ag, c1, i1)) continue;
            int k1 = c1->kind[i1];
            if (hybrid && o->elem_of_kind[k1] < 0) continue;
            for (int jc = 0; jc < c1->cnb2len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb2[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    const opot *p = &o->pot[k1][c2->kind[i2]];
                    if (!hybrid || p->cat == ORC_POT_EAM) eam_rho2(e, p, c1, i1, c2, i2);
                }
            }
            for (int i2 = 0; i2 < i1; i2++)
            {
                if (!ACTIVE(ag, c1, i2)) continue;
                const opot *p = &o->pot[k1][c1->kind[i2]];
                if (!hybrid || p->cat == ORC_POT_EAM) eam_rho2(e, p, c1, i1, c1, i2);
            }
        }
    }
    for (int ic = o->nci; ic < o->ncm; ic++)
    {
        ocell *c1 = &o->grid[ic];
        for (int i1 = 0; i1 < c1->n; i1++)
        {
            if (!ACTIVE(ag, c1, i1)) continue;
            int k1 = c1->kind[i1];
            if (hybrid && o->elem_of_kind[k1] < 0) continue;
            for (int jc = 0; jc < c1->cnb0len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb0[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    const opot *p = &o->pot[k1][c2->kind[i2]];
                    if (!hybrid || p->cat == ORC_POT_EAM) eam_rho1(e, p, c1, i1, c2, i2);
                }
            }
        }
    }
    o->FembSum = calc_femb_prime(o, ag);
}

/* dispatch of one pair in the hybrid pass 0: HYBRID_PASS0_MACRO2/1, forces.c:141-191 */
static inline void hybrid_pair(orc_t *o, int which, ocell *c1, int i1, ocell *c2, int i2, double *pe)
{
    const opot *p = &o->pot[c1->kind[i1]][c2->kind[i2]];
    switch (p->cat)
    {
        case ORC_POT_EAM:
            if (which == 2) eam_force_pe2(&o->eam, p, c1, i1, c2, i2, pe);
            else eam_force_pe1(&o->eam, p, c1, i1, c2, i2, pe);
            break;
        case ORC_POT_LJ: pair_lj(p, c1, i1, c2, i2, pe); break;
        case ORC_POT_MORSE: pair_morse(p, c1, i1, c2, i2, pe); break;
    }
}

/* _fmd_computeEAM_pass0 (hybrid=0) and compute_hybrid_pass0 (hybrid=1),
   eam.c:29-146, forces.c:194-290 */
static void compute_pass0(orc_t *o, int ag, int hybrid)
{
    oeam *e = &o->eam;
    double pe = 0.0;
    clean_forces(o);
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c1 = &o->grid[ic];
        for (int i1 = 0; i1 < c1->n; i1++)
        {
            if (!ACTIVE(ag, c1, i1)) continue;
            for (int jc = 0; jc < c1->cnb2len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb2[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    if (hybrid) hybrid_pair(o, 2, c1, i1, c2, i2, &pe);
                    else eam_force_pe2(e, &o->pot[c1->kind[i1]][c2->kind[i2]], c1, i1, c2, i2, &pe);
                }
            }
            for (int i2 = 0; i2 < i1; i2++)
            {
                if (!ACTIVE(ag, c1, i2)) continue;
                if (hybrid) hybrid_pair(o, 2, c1, i1, c1, i2, &pe);
                else eam_force_pe2(e, &o->pot[c1->kind[i1]][c1->kind[i2]], c1, i1, c1, i2, &pe);
            }
        }
    }
    for (int ic = o->nci; ic < o->ncm; ic++)
    {
        ocell *c1 = &o->grid[ic];
        for (int i1 = 0; i1 < c1->n; i1++)
        {
            if (!ACTIVE(ag, c1, i1)) continue;
            for (int jc = 0; jc < c1->cnb1len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb1[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    const opot *p = &o->pot[c1->kind[i1]][c2->kind[i2]];
                    if (!hybrid || p->cat == ORC_POT_EAM) eam_force(e, p, c1, i1, c2, i2);
                }
            }
            for (int jc = 0; jc < c1->cnb2len; jc++)
            {
                ocell *c2 = &o->grid[c1->cnb2[jc]];
                for (int i2 = 0; i2 < c2->n; i2++)
                {
                    if (!ACTIVE(ag, c2, i2)) continue;
                    if (hybrid) hybrid_pair(o, 1, c1, i1, c2, i2, &pe);
                    else eam_force_pe1(e, &o->pot[c1->kind[i1]][c2->kind[i2]], c1, i1, c2, i2, &pe);
                }
            }
        }
    }
    o->PE = pe + o->FembSum;
}

/* add_ttm_term_to_forces, forces.c:292-330 (1-D only) */
static void ttm_forces(orc_t *o, int ag);

/* _fmd_dync_updateForces, forces.c:332-382 */
void orc_forces(orc_t *o, int ag)
{
    int has[4];
    int n = potcats(o, has);
    ghost_init(o);
    if (n == 1)
    {
        if (has[ORC_POT_LJ]) compute_pairwise(o, ag, ORC_POT_LJ);
        else if (has[ORC_POT_MORSE]) compute_pairwise(o, ag, ORC_POT_MORSE);
        else
        {
            compute_pass1(o, ag, 0);
            ghost_femb(o);
            compute_pass0(o, ag, 0);
        }
    }
    else
    {
        o->FembSum = 0.0;
        if (has[ORC_POT_EAM])
        {
            compute_pass1(o, ag, 1);
            ghost_femb(o);
        }
        compute_pass0(o, ag, 1);
    }
    ghost_fold(o);
    if (o->ttm.on) ttm_forces(o, ag);
    for (int ic = o->nci; ic < o->ncm; ic++) o->grid[ic].n = 0;   /* _fmd_ghostparticles_clean */
}

/* ----------------------------------------------------------- integrators */

/* updateGroupProperties, integrators.c:78-92 */
static void group_props(orc_t *o, const double p[3], int n, double mv2)
{
    for (int d = 0; d < 3; d++) o->p[d] = p[d];
    o->N = n;
    o->KE = 0.5 * mv2;
    o->T = mv2 / (3.0 * n * ORC_KB);
}

/* _fmd_compute_GroupTemperature_etc_localgrid, matter.c:31-68 */
void orc_thermo(orc_t *o, int ag)
{
    double mv2 = 0, p[3] = {0, 0, 0};
    int n = 0;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            n++;
            double m = o->mass[c->kind[i]];
            double *v = &c->v[3 * i];
            for (int d = 0; d < 3; d++) p[d] += m * v[d];
            mv2 += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
    }
    group_props(o, p, n, mv2);
}

/* VelocityVerlet_startStep, integrators.c:40-75 (T from the previous finish) */
void orc_vv_start(orc_t *o, int ag, double dt, int thermostat, double tau, double T0)
{
    double lambda = 1.0;
    if (thermostat) lambda = sqrt(1 + dt / tau * (T0 / o->T - 1));
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            double m = o->mass[c->kind[i]];
            for (int d = 0; d < 3; d++)
            {
                if (thermostat) c->v[3 * i + d] *= lambda;
                c->v[3 * i + d] += dt * 0.5 / m * c->F[3 * i + d];
                c->x[3 * i + d] += dt * c->v[3 * i + d];
            }
        }
    }
    orc_refresh(o);
}

/* VelocityVerlet_finishStep, integrators.c:94-126 */
void orc_vv_finish(orc_t *o, int ag, double dt)
{
    double mv2 = 0, p[3] = {0, 0, 0};
    int n = 0;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            n++;
            double m = o->mass[c->kind[i]];
            double *v = &c->v[3 * i];
            for (int d = 0; d < 3; d++)
            {
                v[d] += dt * 0.5 / m * c->F[3 * i + d];
                p[d] += m * v[d];
            }
            mv2 += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
    }
    group_props(o, p, n, mv2);
}

/* SymplecticEuler_takeOneStep, integrators.c:129-179 */
void orc_se_step(orc_t *o, int ag, double dt)
{
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            double m = o->mass[c->kind[i]];
            for (int d = 0; d < 3; d++)
            {
                c->v[3 * i + d] += dt * c->F[3 * i + d] / m;
                c->x[3 * i + d] += dt * c->v[3 * i + d];
            }
        }
    }
    orc_refresh(o);
    double mv2 = 0, p[3] = {0, 0, 0};
    int n = 0;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            n++;
            double m = o->mass[c->kind[i]];
            double t = 0.0;
            for (int d = 0; d < 3; d++)
            {
                t += c->v[3 * i + d] * c->v[3 * i + d];
                p[d] += m * c->v[3 * i + d];
            }
            mv2 += m * t;
        }
    }
    group_props(o, p, n, mv2);
}

void orc_get_energies(orc_t *o, double *out)
{
    out[0] = o->PE; out[1] = o->KE; out[2] = o->T;
    out[3] = o->p[0]; out[4] = o->p[1]; out[5] = o->p[2];
    out[6] = o->N; out[7] = o->FembSum;
}

int orc_error_count(orc_t *o, int *last)
{
    if (last) *last = o->lasterr;
    return o->nerr;
}

/* _fmd_timer_is_its_time, timer.c:27-36 */
int orc_timer_is_its_time(double t, double dt_half, double start, double interval)
{
    double tc = round((t - start) / interval) * interval + start;
    double d = t - tc;
    return (d >= -dt_half && d < dt_half);
}

/* VelocityVerlet_integrate, integrators.c:227-264 (no turies, no events) */
int orc_integrate_vv(orc_t *o, int ag, double *time, double duration, double dt,
                     int thermostat, double tau, double T0)
{
    double start = *time;
    int steps = 0;
    orc_forces(o, ag);
    if (thermostat) orc_thermo(o, ag);   /* KineticEnergyUpdated was false, integrators.c:46 */
    while (*time < start + duration)
    {
        orc_vv_start(o, ag, dt, thermostat, tau, T0);
        orc_forces(o, ag);
        orc_vv_finish(o, ag, dt);
        *time += dt;
        steps++;
    }
    return steps;
}

/* ------------------------------------------------------------------- TTM */

/* turi geometry, turi.c:648-738 with ns=1: tcellh = l / tdims.  The
   reference's turi is 1 x 1 x ncz; ncx x ncy x ncz is this build's extension. */
int orc_ttm_init(orc_t *o, const orc_ttm_cfg_t *cfg, double Te_init)
{
    ottm *t = &o->ttm;
    t->c = *cfg;
    t->ncx = cfg->ncx > 0 ? cfg->ncx : 1;
    t->ncy = cfg->ncy > 0 ? cfg->ncy : 1;
    t->nzg = cfg->ncz + 2;
    t->n = t->ncx * t->ncy * t->nzg;
    t->dim3 = t->ncx * t->ncy > 1 || cfg->kmode != 0 || cfg->solver3d;
    t->tcellh[0] = o->l[0] / t->ncx;
    t->tcellh[1] = o->l[1] / t->ncy;
    t->tcellh[2] = o->l[2] / cfg->ncz;
    t->tvol = t->tcellh[0] * t->tcellh[1] * t->tcellh[2];
    t->num = calloc(t->n, sizeof(unsigned));
    t->mass = calloc(t->n, sizeof(double));
    t->vcm = calloc(3 * t->n, sizeof(double));
    t->Ti = calloc(t->n, sizeof(double));
    t->Te = calloc(t->n, sizeof(double));
    t->Te2 = calloc(t->n, sizeof(double));   /* ghosts: the reference leaves them uninitialised */
    t->xi = calloc(t->n, sizeof(double));
    t->Ke = calloc(t->n, sizeof(double));
    t->Ce = calloc(t->n, sizeof(double));
    t->Gc = calloc(t->n, sizeof(double));
    /* fmd_ttm_setElectronTemperature, ttm.c:739-755: interior cells */
    for (int a = 0; a < t->ncx * t->ncy; a++)
        for (int i = 1; i <= cfg->ncz; i++) t->Te[a * t->nzg + i] = Te_init;
    if (cfg->type == 2)
    {
        t->T_C = malloc(cfg->nC * sizeof(double)); t->Ct = malloc(cfg->nC * sizeof(double));
        t->T_G = malloc(cfg->nG * sizeof(double)); t->Gt = malloc(cfg->nG * sizeof(double));
        memcpy(t->T_C, cfg->T_C, cfg->nC * sizeof(double)); memcpy(t->Ct, cfg->Ct, cfg->nC * sizeof(double));
        memcpy(t->T_G, cfg->T_G, cfg->nG * sizeof(double)); memcpy(t->Gt, cfg->Gt, cfg->nG * sizeof(double));
        t->CDD = malloc(cfg->nC * sizeof(double)); t->GDD = malloc(cfg->nG * sizeof(double));
        orc_spline_prepare_adv(t->T_C, t->Ct, cfg->nC, t->CDD);
        orc_spline_prepare_adv(t->T_G, t->Gt, cfg->nG, t->GDD);
    }
    t->on = 1;
    return 0;
}

/* fmd_matt_findLimits (matter.c:538-566) + _fmd_ttm_getReady (ttm.c:63-80) */
void orc_ttm_get_ready(orc_t *o, int ag)
{
    ottm *t = &o->ttm;
    double L[3], U[3];
    for (int d = 0; d < 3; d++) { L[d] = DBL_MAX; U[d] = DBL_MIN; }
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
            if (ag == -1 || c->grp[i] == ag)
                for (int d = 0; d < 3; d++)
                {
                    if (c->x[3 * i + d] < L[d]) L[d] = c->x[3 * i + d];
                    if (c->x[3 * i + d] > U[d]) U[d] = c->x[3 * i + d];
                }
    }
    t->frontsurf = L[2];
    double vol = 1.0;
    for (int d = 0; d < 3; d++) vol *= U[d] - L[d];
    t->initial_atoms = o->N / vol * t->tvol;
    t->min_atoms = t->initial_atoms * t->c.act_frac;
}

static inline int ttm_iz(orc_t *o, const double *x)
{
    return (int)(x[2] / o->ttm.tcellh[2]) + 1;
}

/* update_field_vcm_and_mass + update_field_temperature_and_number, turi.c:1112-1261.
   These sums run over every atom of the subdomain (no group filter). */
static void ttm_fields(orc_t *o)
{
    ottm *t = &o->ttm;
    for (int i = 1; i <= t->c.ncz; i++)
    {
        t->mass[i] = 0.0;
        t->vcm[3 * i] = t->vcm[3 * i + 1] = t->vcm[3 * i + 2] = 0.0;
    }
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            int iz = ttm_iz(o, &c->x[3 * i]);
            double m = o->mass[c->kind[i]];
            t->mass[iz] += m;
            for (int d = 0; d < 3; d++) t->vcm[3 * iz + d] += m * c->v[3 * i + d];
        }
    }
    for (int i = 1; i <= t->c.ncz; i++)
        for (int d = 0; d < 3; d++) t->vcm[3 * i + d] /= t->mass[i];
    for (int i = 1; i <= t->c.ncz; i++) { t->Ti[i] = 0.0; t->num[i] = 0; }
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            int iz = ttm_iz(o, &c->x[3 * i]);
            double vt[3];
            for (int d = 0; d < 3; d++) vt[d] = c->v[3 * i + d] - t->vcm[3 * iz + d];
            double m = o->mass[c->kind[i]];
            t->Ti[iz] += m * (vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
            ++t->num[iz];
        }
    }
    for (int i = 1; i <= t->c.ncz; i++) t->Ti[i] /= (3.0 * ORC_KB * t->num[i]);
}

/* FIND_klo_AND_ETC, cspline.h:29-44 */
static void find_klo(const double *list, double x, int n, int *klo, int *khi, double *h, double *a, double *b)
{
    int lo = 0, hi = n - 1, k;
    while (hi - lo > 1)
    {
        k = (hi + lo) >> 1;
        if (list[k] > x) hi = k;
        else lo = k;
    }
    *h = list[hi] - list[lo];
    *a = (list[hi] - x) / *h;
    *b = (x - list[lo]) / *h;
    *klo = lo; *khi = hi;
}

/* type2_update_Ce_Ke_G, ttm.c:154-192 */
static void ttm_type2_update(orc_t *o)
{
    ottm *t = &o->ttm;
    for (int i = 1; i <= t->c.ncz; i++)
    {
        if (t->Te[i] < 0.) continue;
        if (t->Te[i] < t->T_C[0] || t->Te[i] > t->T_C[t->c.nC - 1]) orc_err(o, 5);
        int klo, khi;
        double h, a, b;
        find_klo(t->T_C, t->Te[i], t->c.nC, &klo, &khi, &h, &a, &b);
        t->Ce[i] = spl_val(a, b, t->Ct, klo, khi, t->CDD, h);
        t->Ke[i] = 1. / 3. * t->Ce[i] * t->c.Kzh_v2 / (t->c.Kzh_A * (t->Te[i] * t->Te[i]) + t->c.Kzh_B * t->Ti[i]);
        if (t->Te[i] < t->T_G[0] || t->Te[i] > t->T_G[t->c.nG - 1]) orc_err(o, 5);
        find_klo(t->T_G, t->Te[i], t->c.nG, &klo, &khi, &h, &a, &b);
        t->Gc[i] = spl_val(a, b, t->Gt, klo, khi, t->GDD, h);
    }
}

/* presolve_common_types_1_2, ttm.c:194-230 */
static void ttm_presolve(orc_t *o)
{
    ottm *t = &o->ttm;
    for (int i = 1; i <= t->c.ncz; i++)
    {
        t->xi[i] = 0.0;
        if (t->Te[i] < 0.0 && t->num[i] >= t->min_atoms)
        {
            int count = 0;
            double sum = 0.0;
            int im1 = (i == 1) ? 0 : i - 1;
            if (t->Te[im1] >= 0.0) { sum += t->Te[im1] * t->num[im1]; count += t->num[im1]; }
            if (t->Te[i + 1] >= 0.0) { sum += t->Te[i + 1] * t->num[i + 1]; count += t->num[i + 1]; }
            t->Te[i] = (count > 0) ? sum / count : t->Ti[i];
        }
        if (t->num[i] < t->min_atoms) t->Te[i] = t->Te2[i] = -1.0;
    }
}

/* ttm_type1_solve_1d / ttm_type2_solve_1d, ttm.c:250-383 */
static void ttm_solve(orc_t *o, double time, double mddt)
{
    ottm *t = &o->ttm;
    int CalcSource = fabs(time - t->c.laser_t0) < t->c.laser_tdiff;
    double dt = mddt / t->c.ratio;
    for (int j = 0; j < t->c.ratio; j++)
    {
        double spc = 0.0;
        if (CalcSource)
        {
            double q = time + j * dt - t->c.laser_t0;
            spc = t->c.laser_factor * exp(t->c.laser_2sig2_inv * (q * q));
        }
        for (int i = 1; i <= t->c.ncz; i++)
        {
            if (t->Te[i] < 0) continue;
            double source = 0.0;
            if (CalcSource)
            {
                double z = (i - 1 + 0.5) * t->tcellh[2];
                source = spc * exp(t->c.laser_absdepth_inv * (z - t->frontsurf));
            }
            int im1 = (i == 1) ? 0 : i - 1;
            int ilo = (t->Te[im1] < 0) ? i : im1;
            int ihi = (t->Te[i + 1] < 0) ? i : i + 1;
            if (t->c.type == 1)
                t->Te2[i] = t->Te[i] + dt / (t->c.C_gamma * t->Te[i]) * (
                    t->c.K * (t->Te[ihi] - 2 * t->Te[i] + t->Te[ilo]) / (t->tcellh[2] * t->tcellh[2])
                    - t->c.G * (t->Te[i] - t->Ti[i]) + source);
            else
            {
                double KeL = 0.5 * (t->Ke[i] + t->Ke[ilo]);
                double KeR = 0.5 * (t->Ke[i] + t->Ke[ihi]);
                t->Te2[i] = t->Te[i] + dt / t->Ce[i] * (
                    (KeR * (t->Te[ihi] - t->Te[i]) - KeL * (t->Te[i] - t->Te[ilo])) / (t->tcellh[2] * t->tcellh[2])
                    - t->Gc[i] * (t->Te[i] - t->Ti[i]) + source);
            }
            t->xi[i] += t->Te2[i];
        }
        double *tmp = t->Te2; t->Te2 = t->Te; t->Te = tmp;
        if (t->c.type == 2 && j < t->c.ratio - 1) ttm_type2_update(o);
    }
    for (int i = 1; i <= t->c.ncz; i++)
        if (t->num[i] >= t->min_atoms)
        {
            double G = t->c.type == 1 ? t->c.G : t->Gc[i];
            t->xi[i] = G * t->tvol * (t->xi[i] / t->c.ratio - t->Ti[i]) / (3 * ORC_KB * t->num[i] * t->Ti[i]);
        }
}

/* ---- ncx x ncy x ncz turies (this build's extension; periodic in x and y,
   a ghost layer in z as in the 1-D turi) ---- */

static inline int tidx(const ottm *t, int ix, int iy, int iz) { return (ix * t->ncy + iy) * t->nzg + iz; }

/* turi cell of a position: (int)(x / tcellh) per dimension, z shifted by the
   ghost layer (turi.c:1136 with itc_glob_to_loc); x and y are clamped to the
   grid, where a coordinate within rounding of the box length lands */
static int ttm_cell3(const ottm *t, const double *x)
{
    double fx = x[0] / t->tcellh[0], fy = x[1] / t->tcellh[1], fz = x[2] / t->tcellh[2];
    int ix = fx >= 0.0 ? (fx < t->ncx ? (int)fx : t->ncx - 1) : 0;
    int iy = fy >= 0.0 ? (fy < t->ncy ? (int)fy : t->ncy - 1) : 0;
    /* a non-finite or far-away z goes to the lower ghost cell */
    int iz = (fz >= -1.0 && fz < t->nzg - 1) ? (int)fz + 1 : 0;
    return tidx(t, ix, iy, iz);
}

static void ttm_fields3(orc_t *o)
{
    ottm *t = &o->ttm;
    for (int a = 0; a < t->n; a++)
    {
        t->mass[a] = 0.0;
        t->vcm[3 * a] = t->vcm[3 * a + 1] = t->vcm[3 * a + 2] = 0.0;
        t->Ti[a] = 0.0;
        t->num[a] = 0;
    }
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            int a = ttm_cell3(t, &c->x[3 * i]);
            double m = o->mass[c->kind[i]];
            t->mass[a] += m;
            for (int d = 0; d < 3; d++) t->vcm[3 * a + d] += m * c->v[3 * i + d];
        }
    }
    for (int a = 0; a < t->n; a++)
        if (t->mass[a] != 0.0)
            for (int d = 0; d < 3; d++) t->vcm[3 * a + d] /= t->mass[a];
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            int a = ttm_cell3(t, &c->x[3 * i]);
            double vt[3];
            for (int d = 0; d < 3; d++) vt[d] = c->v[3 * i + d] - t->vcm[3 * a + d];
            double m = o->mass[c->kind[i]];
            t->Ti[a] += m * (vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
            ++t->num[a];
        }
    }
    for (int a = 0; a < t->n; a++)
        if (t->num[a]) t->Ti[a] /= (3.0 * ORC_KB * t->num[a]);
}

/* Activation / deactivation on the ncx x ncy x ncz turi (this build's
   extension of presolve_common_types_1_2, ttm.c:194-230).  Each (x, y) column
   is scanned upwards in z exactly as the reference's 1-D scan: the cell below
   has already been processed in this pass, the cell above has not.  The x and
   y neighbours contribute their Te from before the pass, so the columns are
   independent and a 1 x 1 x n turi reduces to the reference scan (its x and y
   neighbours are the cell itself, inactive while it is being activated).  A
   newly active cell takes the atom-weighted mean Te of its active neighbours
   in the order z-, z+, x-, x+, y-, y+, or its Ti when it has none. */
static void ttm_presolve3(orc_t *o)
{
    ottm *t = &o->ttm;
    double *before = malloc(t->n * sizeof(double));
    memcpy(before, t->Te, t->n * sizeof(double));
    for (int ix = 0; ix < t->ncx; ix++)
        for (int iy = 0; iy < t->ncy; iy++)
            for (int iz = 1; iz <= t->c.ncz; iz++)
            {
                int a = tidx(t, ix, iy, iz);
                t->xi[a] = 0.0;
                if (t->Te[a] < 0.0 && t->num[a] >= t->min_atoms)
                {
                    const int nb[6] = {tidx(t, ix, iy, iz - 1), tidx(t, ix, iy, iz + 1),
                                       tidx(t, (ix + t->ncx - 1) % t->ncx, iy, iz), tidx(t, (ix + 1) % t->ncx, iy, iz),
                                       tidx(t, ix, (iy + t->ncy - 1) % t->ncy, iz), tidx(t, ix, (iy + 1) % t->ncy, iz)};
                    int count = 0;
                    double sum = 0.0;
                    for (int k = 0; k < 6; k++)
                    {
                        const double te = k < 2 ? t->Te[nb[k]] : before[nb[k]];
                        if (te >= 0.0) { sum += te * t->num[nb[k]]; count += t->num[nb[k]]; }
                    }
                    t->Te[a] = (count > 0) ? sum / count : t->Ti[a];
                }
                if (t->num[a] < t->min_atoms) t->Te[a] = t->Te2[a] = -1.0;
            }
    free(before);
}

/* explicit FD sub-steps on the ncx x ncy x ncz turi.  Ce = C_gamma * Te, G
   constant; kmode 0: K (Te_hi - 2 Te + Te_lo) / h^2 per dimension, the 1-D
   type-1 expression (ttm.c:281-287); kmode 1: kappa = K * Te / Ti with face
   values (kappa_i + kappa_nb) / 2, the type-2 flux form (ttm.c:347-352).
   An inactive neighbour is replaced by the cell itself (zero flux). */
static void ttm_solve3(orc_t *o, double time, double mddt)
{
    ottm *t = &o->ttm;
    int CalcSource = fabs(time - t->c.laser_t0) < t->c.laser_tdiff;
    double dt = mddt / t->c.ratio;
    const double h2[3] = {t->tcellh[0] * t->tcellh[0], t->tcellh[1] * t->tcellh[1], t->tcellh[2] * t->tcellh[2]};
    for (int j = 0; j < t->c.ratio; j++)
    {
        double spc = 0.0;
        if (CalcSource)
        {
            double q = time + j * dt - t->c.laser_t0;
            spc = t->c.laser_factor * exp(t->c.laser_2sig2_inv * (q * q));
        }
        for (int ix = 0; ix < t->ncx; ix++)
            for (int iy = 0; iy < t->ncy; iy++)
                for (int iz = 1; iz <= t->c.ncz; iz++)
                {
                    int a = tidx(t, ix, iy, iz);
                    if (t->Te[a] < 0) continue;
                    double source = 0.0;
                    if (CalcSource)
                    {
                        double z = (iz - 1 + 0.5) * t->tcellh[2];
                        source = spc * exp(t->c.laser_absdepth_inv * (z - t->frontsurf));
                    }
                    int lo[3] = {tidx(t, (ix + t->ncx - 1) % t->ncx, iy, iz), tidx(t, ix, (iy + t->ncy - 1) % t->ncy, iz),
                                 tidx(t, ix, iy, iz - 1)};
                    int hi[3] = {tidx(t, (ix + 1) % t->ncx, iy, iz), tidx(t, ix, (iy + 1) % t->ncy, iz),
                                 tidx(t, ix, iy, iz + 1)};
                    for (int d = 0; d < 3; d++)
                    {
                        if (t->Te[lo[d]] < 0) lo[d] = a;
                        if (t->Te[hi[d]] < 0) hi[d] = a;
                    }
                    double flux = 0.0;
                    if (t->c.kmode == 0)
                    {
                        for (int d = 0; d < 3; d++)
                            flux += t->c.K * (t->Te[hi[d]] - 2 * t->Te[a] + t->Te[lo[d]]) / h2[d];
                    }
                    else
                    {
                        double ka = t->c.K * t->Te[a] / t->Ti[a];
                        for (int d = 0; d < 3; d++)
                        {
                            double KeL = 0.5 * (ka + t->c.K * t->Te[lo[d]] / t->Ti[lo[d]]);
                            double KeR = 0.5 * (ka + t->c.K * t->Te[hi[d]] / t->Ti[hi[d]]);
                            flux += (KeR * (t->Te[hi[d]] - t->Te[a]) - KeL * (t->Te[a] - t->Te[lo[d]])) / h2[d];
                        }
                    }
                    t->Te2[a] = t->Te[a] + dt / (t->c.C_gamma * t->Te[a]) * (flux - t->c.G * (t->Te[a] - t->Ti[a]) +
                                                                               source);
                    t->xi[a] += t->Te2[a];
                }
        double *tmp = t->Te2; t->Te2 = t->Te; t->Te = tmp;
    }
    for (int a = 0; a < t->n; a++)
    {
        int iz = a % t->nzg;
        if (iz >= 1 && iz <= t->c.ncz && t->num[a] >= t->min_atoms)
            t->xi[a] = t->c.G * t->tvol * (t->xi[a] / t->c.ratio - t->Ti[a]) / (3 * ORC_KB * t->num[a] * t->Ti[a]);
    }
}

void orc_ttm_set_fields(orc_t *o, const double *Te, const double *Ti, const unsigned *num, double frontsurf,
                        unsigned min_atoms)
{
    ottm *t = &o->ttm;
    for (int a = 0; a < t->n; a++)
    {
        t->Te[a] = Te[a];
        t->Te2[a] = Te[a];
        t->Ti[a] = Ti[a];
        t->num[a] = num[a];
        t->xi[a] = 0.0;
    }
    t->frontsurf = frontsurf;
    t->min_atoms = min_atoms;
}

void orc_ttm_solve_only(orc_t *o, double time, double dt)
{
    ottm *t = &o->ttm;
    if (t->dim3) ttm_solve3(o, time, dt);
    else
    {
        for (int i = 1; i <= t->c.ncz; i++) t->xi[i] = 0.0;
        if (t->c.type == 2) ttm_type2_update(o);
        ttm_solve(o, time, dt);
    }
}

/* presolve (activation / deactivation, ttm.c:194-248) and solve on injected fields */
void orc_ttm_presolve_solve(orc_t *o, double time, double dt)
{
    ottm *t = &o->ttm;
    if (t->dim3)
    {
        ttm_presolve3(o);
        ttm_solve3(o, time, dt);
        return;
    }
    ttm_presolve(o);
    if (t->c.type == 2) ttm_type2_update(o);
    ttm_solve(o, time, dt);
}

void orc_ttm_update(orc_t *o, double time, double dt)
{
    ottm *t = &o->ttm;
    if (!t->on) return;
    if (!((time >= t->c.start - dt / 2.0) && ((time < t->c.stop + dt / 2.0) || t->c.stop < t->c.start)))
        return;
    if (t->dim3)
    {
        ttm_fields3(o);
        ttm_presolve3(o);
        ttm_solve3(o, time, dt);
        return;
    }
    ttm_fields(o);
    ttm_presolve(o);
    if (t->c.type == 2) ttm_type2_update(o);
    ttm_solve(o, time, dt);
}

static double g_ttm_time, g_ttm_dt;

/* add_ttm_term_to_forces, forces.c:292-330 */
static void ttm_forces(orc_t *o, int ag)
{
    ottm *t = &o->ttm;
    double time = g_ttm_time, dt = g_ttm_dt;
    if (!((time >= t->c.start - dt / 2.0) && ((time < t->c.stop + dt / 2.0) || t->c.stop < t->c.start)))
        return;
    for (int ic = 0; ic < o->nci; ic++)
    {
        ocell *c = &o->grid[ic];
        for (int i = 0; i < c->n; i++)
        {
            if (!ACTIVE(ag, c, i)) continue;
            double m = o->mass[c->kind[i]];
            if (t->ncx * t->ncy > 1)
            {
                /* forces.c:317-328, a turi that is not 1 x 1 x n: every component */
                int a = ttm_cell3(t, &c->x[3 * i]);
                double ximass = t->xi[a] * m;
                for (int d = 0; d < 3; d++) c->F[3 * i + d] += ximass * (c->v[3 * i + d] - t->vcm[3 * a + d]);
                continue;
            }
            int iz = ttm_iz(o, &c->x[3 * i]);
            c->F[3 * i + 2] += t->xi[iz] * m * (c->v[3 * i + 2] - t->vcm[3 * iz + 2]);
        }
    }
}

void orc_ttm_arrays(orc_t *o, double *Te, double *Ti, double *xi, double *vcmz, unsigned *num, double *mass)
{
    ottm *t = &o->ttm;
    for (int i = 0; i < t->n; i++)
    {
        if (Te) Te[i] = t->Te[i];
        if (Ti) Ti[i] = t->Ti[i];
        if (xi) xi[i] = t->xi[i];
        if (vcmz) vcmz[i] = t->vcm[3 * i + 2];
        if (num) num[i] = t->num[i];
        if (mass) mass[i] = t->mass[i];
    }
}

void orc_ttm_params(orc_t *o, double *frontsurf, unsigned *min_atoms, double *tcellh_z)
{
    *frontsurf = o->ttm.frontsurf;
    *min_atoms = o->ttm.min_atoms;
    *tcellh_z = o->ttm.tcellh[2];
}

/* SymplecticEuler_integrate, integrators.c:181-225 (no events) */
int orc_integrate_se(orc_t *o, int ag, double *time, double duration, double dt)
{
    double start = *time;
    int steps = 0;
    g_ttm_dt = dt;
    if (o->ttm.on) orc_ttm_get_ready(o, ag);
    g_ttm_time = *time;
    orc_ttm_update(o, *time, dt);
    orc_forces(o, ag);
    while (*time < start + duration)
    {
        orc_se_step(o, ag, dt);
        *time += dt;
        steps++;
        g_ttm_time = *time;
        orc_ttm_update(o, *time, dt);
        orc_forces(o, ag);
    }
    return steps;
}
