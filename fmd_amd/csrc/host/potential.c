/*
 * potential.c -- atom kinds, the pair-potential table, the setfl EAM loader
 * and the natural cubic splines the EAM tables are built with.
 *
 * Reference behaviour followed:
 *   fmd_matt_setAtomKinds          potential.c:31-54 (mass -> amu / 9648.5332907)
 *   fmd_pot_apply / pottable       potential.c:122-175 (symmetric, iloc/jloc)
 *   category scan                  potential.c:177-249
 *   largest cutoff                 potential.c:251-291
 *   fmd_pot_lj_apply / morse       lj.c:92-113, morse.c:94-114
 *   setfl loader + r -> r^2 tables eam.c:324-555: Nr += 2 zero-padded,
 *                                  r*phi(0) replaced by FLT_MAX, each table
 *                                  splined in r and resampled at sqrt(i*dr2),
 *                                  then splined again on the r^2 grid
 *   splines                        cspline.c:24-117 (natural boundary)
 * The tables produced here are the exact arrays the reference builds; the
 * device layer turns them into per-interval polynomial coefficients.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_host.h"

/* ----------------------------------------------------------------- spline */

void fmdh_spline_prepare(double h, const double *y, int n, double *yDD)
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

void fmdh_spline_prepare_adv(const double *x, const double *y, int n, double *yDD)
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

double fmdh_spline_val(double h, const double *ya, const double *yDDa, double x)
{
    int lo = (int)(x / h), hi = lo + 1;
    double a = hi - x / h, b = 1 - a;
    return (a * ya[lo] + b * ya[hi] + ((a * a * a - a) * yDDa[lo] + (b * b * b - b) * yDDa[hi]) * (h * h) / 6.0);
}

/* ------------------------------------------------------------------ kinds */

void fmd_matt_setAtomKinds(fmd_t *md, unsigned number, const fmd_string_t names[], const fmd_real_t masses[])
{
    if (number == 0 || number > HMAXK)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "number of atomkinds", (int)number);
        return;
    }
    md->nkinds = number;
    md->kinds = calloc(number, sizeof(atomkind_t));
    for (unsigned i = 0; i < number; i++)
    {
        md->kinds[i].mass = masses[i] / MD_MASS_UNIT;
        md->kinds[i].name = strdup(names[i]);
        md->kinds[i].eam_element = -1;
    }
}

static void pottable_create(fmd_t *md)
{
    md->pottable = calloc((size_t)md->nkinds * md->nkinds, sizeof(potpair_t));
}

static void potlist_add(fmd_t *md, struct _fmd_pot *p)
{
    struct _fmd_pot **l = realloc(md->potlist, (md->npots + 1) * sizeof(*l));
    if (!l) return;
    md->potlist = l;
    md->potlist[md->npots++] = p;
}

static int eam_find(const eam_t *e, const char *name)
{
    for (int i = 0; i < e->nelem; i++)
        if (strcmp(name, e->el[i].name) == 0) return i;
    return -1;
}

void fmd_pot_apply(fmd_t *md, unsigned atomkind1, unsigned atomkind2, fmd_pot_t *pot)
{
    if (!pot || atomkind1 >= md->nkinds || atomkind2 >= md->nkinds)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "atomkind", (int)(atomkind1 > atomkind2 ? atomkind1 : atomkind2));
        return;
    }
    if (md->pottable == NULL) pottable_create(md);
    unsigned n = md->nkinds;
    if (pot->cat == POT_EAM_ALLOY)
    {
        eam_t *e = pot->data;
        int l1 = eam_find(e, md->kinds[atomkind1].name);
        int l2 = eam_find(e, md->kinds[atomkind2].name);
        if (l1 < 0 || l2 < 0)
        {
            fmdh_err_wrong_potential(md, false, HERE, "EAM", atomkind1, atomkind2);
            return;
        }
        md->pottable[atomkind1 * n + atomkind2].iloc = l1;
        md->pottable[atomkind1 * n + atomkind2].jloc = l2;
        md->pottable[atomkind2 * n + atomkind1].iloc = l2;
        md->pottable[atomkind2 * n + atomkind1].jloc = l1;
        md->kinds[atomkind1].eam_element = l1;
        md->kinds[atomkind2].eam_element = l2;
        md->eam = e;
    }
    md->pottable[atomkind1 * n + atomkind2].cat = md->pottable[atomkind2 * n + atomkind1].cat = pot->cat;
    md->pottable[atomkind1 * n + atomkind2].data = md->pottable[atomkind2 * n + atomkind1].data = pot->data;
    md->potcats_valid = false;
    md->dev_pot_dirty = true;
}

fmd_pot_t *fmd_pot_lj_apply(fmd_t *md, unsigned atomkind1, unsigned atomkind2, fmd_real_t sigma,
                            fmd_real_t epsilon, fmd_real_t cutoff)
{
    lj_t *lj = malloc(sizeof *lj);
    lj->sig = sigma;
    lj->sig_sqr = sigma * sigma;
    lj->eps = epsilon;
    lj->eps4 = 4. * epsilon;
    lj->cutoff_sqr = cutoff * cutoff;
    struct _fmd_pot *pot = malloc(sizeof *pot);
    pot->cat = POT_LJ_6_12;
    pot->data = lj;
    potlist_add(md, pot);
    fmd_pot_apply(md, atomkind1, atomkind2, pot);
    return pot;
}

fmd_pot_t *fmd_pot_morse_apply(fmd_t *md, unsigned atomkind1, unsigned atomkind2, fmd_real_t D0, fmd_real_t alpha,
                               fmd_real_t r0, fmd_real_t cutoff)
{
    morse_t *m = malloc(sizeof *m);
    m->D0 = D0;
    m->alpha = alpha;
    m->r0 = r0;
    m->cutoff_sqr = cutoff * cutoff;
    struct _fmd_pot *pot = malloc(sizeof *pot);
    pot->cat = POT_MORSE;
    pot->data = m;
    potlist_add(md, pot);
    fmd_pot_apply(md, atomkind1, atomkind2, pot);
    return pot;
}

/* ------------------------------------------------------------------- EAM */

/* EAM_convert_r_to_r2, eam.c:324-341.  The resampling reaches the last knot, r = (Nr-1) dr,
   where the spline evaluation (cspline.c:76-89, fmdh_spline_val) takes the interval above it and
   reads one knot past the table with weight zero.  src and its second derivatives therefore carry
   one zero entry past the table (load_setfl allocates src with Nr + 1 zeroed entries): the read
   stays in bounds and adds an exact zero.  Whatever followed an unpadded array could be a NaN or
   an infinity, and 0 x NaN in the last r^2 knot would spread through the natural spline's
   solve to every second derivative of the table. */
static void r_to_r2(const eam_t *e, const double *src, double *dst)
{
    double *dd = calloc((size_t)e->Nr + 1, sizeof(double));
    fmdh_spline_prepare(e->dr, src, e->Nr, dd);
    for (int i = 0; i < e->Nr2; i++) dst[i] = fmdh_spline_val(e->dr, src, dd, sqrt(i * e->dr2));
    free(dd);
}

static void eam_free(eam_t *e)
{
    if (!e) return;
    for (int i = 0; i < e->nelem; i++)
    {
        free(e->el[i].F); free(e->el[i].Fdd); free(e->el[i].rho); free(e->el[i].rhodd);
        for (int j = 0; j <= i; j++) { free(e->phi[i][j]); free(e->phidd[i][j]); }
    }
    free(e);
}

/* load_DYNAMOsetfl, eam.c:376-555 */
static eam_t *load_setfl(fmd_t *md, const char *path)
{
    static const char *ftype = "DYNAMO multi-element setfl (eam/alloy)";
    FILE *fp = fopen(path, "r");
    if (!fp)
    {
        fmdh_err_unable_open_file(md, false, HERE, path);
        return NULL;
    }
    eam_t *e = calloc(1, sizeof *e);
    char str[1024];
    double *tmp = NULL;
    for (int i = 0; i < 3; i++)
        if (!fgets(str, sizeof str, fp)) goto corrupt;
    if (fscanf(fp, "%d", &e->nelem) != 1 || e->nelem < 1 || e->nelem > HMAXE) goto corrupt;
    for (int i = 0; i < e->nelem; i++)
        if (fscanf(fp, "%63s", e->el[i].name) != 1) goto corrupt;
    double cutoff;
    if (fscanf(fp, "%d%lf%d%lf%lf", &e->Nrho, &e->drho, &e->Nr, &e->dr, &cutoff) != 5) goto corrupt;
    if (e->Nrho < 3 || e->Nr < 3) goto corrupt;
    e->Nr2 = (e->Nr += 2);
    if ((e->Nr - 1) * e->dr <= cutoff) goto corrupt;
    e->cutoff_sqr = cutoff * cutoff;
    e->dr2 = ((e->Nr - 1) * e->dr) * ((e->Nr - 1) * e->dr) / (e->Nr2 - 1);
    tmp = calloc((size_t)e->Nr + 1, sizeof(double));   /* one zero past the table: see r_to_r2 */
    for (int i = 0; i < e->nelem; i++)
    {
        eam_element_t *el = &e->el[i];
        if (fscanf(fp, "%1023s%lf%lf%1023s", str, &el->mass, &el->lattice_a, str) != 4) goto corrupt;
        el->mass /= MD_MASS_UNIT;
        el->F = malloc(e->Nrho * sizeof(double));
        for (int j = 0; j < e->Nrho; j++)
            if (fscanf(fp, "%lf", &el->F[j]) != 1) goto corrupt;
        el->rho = malloc(e->Nr2 * sizeof(double));
        for (int j = 0; j < e->Nr - 2; j++)
            if (fscanf(fp, "%lf", &tmp[j]) != 1) goto corrupt;
        tmp[e->Nr - 1] = tmp[e->Nr - 2] = 0.;
        r_to_r2(e, tmp, el->rho);
        el->Fdd = malloc(e->Nrho * sizeof(double));
        el->rhodd = malloc(e->Nr2 * sizeof(double));
        fmdh_spline_prepare(e->drho, el->F, e->Nrho, el->Fdd);
        fmdh_spline_prepare(e->dr2, el->rho, e->Nr2, el->rhodd);
    }
    for (int i = 0; i < e->nelem; i++)
        for (int j = 0; j <= i; j++)
        {
            for (int k = 0; k < e->Nr - 2; k++)
            {
                if (fscanf(fp, "%lf", &tmp[k]) != 1) goto corrupt;
                if (k == 0) tmp[k] = FLT_MAX;
                else tmp[k] /= k * e->dr;
            }
            tmp[e->Nr - 1] = tmp[e->Nr - 2] = 0.;
            e->phi[i][j] = malloc(e->Nr2 * sizeof(double));
            r_to_r2(e, tmp, e->phi[i][j]);
            e->phi[j][i] = e->phi[i][j];
            e->phidd[i][j] = malloc(e->Nr2 * sizeof(double));
            fmdh_spline_prepare(e->dr2, e->phi[i][j], e->Nr2, e->phidd[i][j]);
            e->phidd[j][i] = e->phidd[i][j];
        }
    free(tmp);
    fclose(fp);
    return e;
corrupt:
    free(tmp);
    fclose(fp);
    fmdh_err_file_corrupted(md, true, HERE, ftype, path);
    eam_free(e);
    return NULL;
}

fmd_pot_t *fmd_pot_eam_alloy_load(fmd_t *md, fmd_string_t path)
{
    eam_t *e = load_setfl(md, path);
    if (!e) return NULL;
    struct _fmd_pot *pot = malloc(sizeof *pot);
    pot->cat = POT_EAM_ALLOY;
    pot->data = e;
    potlist_add(md, pot);
    return pot;
}

fmd_real_t fmd_pot_eam_getCutoffRadius(fmd_t *md, fmd_pot_t *pot)
{
    if (pot->cat != POT_EAM_ALLOY)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "potential category", pot->cat);
        return 0.;
    }
    return sqrt(((eam_t *)pot->data)->cutoff_sqr);
}

fmd_real_t fmd_pot_eam_getLatticeParameter(fmd_t *md, fmd_pot_t *pot, fmd_string_t element)
{
    if (pot->cat != POT_EAM_ALLOY)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "potential category", pot->cat);
        return 0.;
    }
    eam_t *e = pot->data;
    int i = eam_find(e, element);
    if (i < 0)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "EAM element", -1);
        return 0.;
    }
    return e->el[i].lattice_a;
}

/* ------------------------------------------------------ categories, cutoff */

static double pair_cutoff_sqr(const potpair_t *p)
{
    switch (p->cat)
    {
        case POT_LJ_6_12: return ((lj_t *)p->data)->cutoff_sqr;
        case POT_MORSE: return ((morse_t *)p->data)->cutoff_sqr;
        case POT_EAM_ALLOY: return ((eam_t *)p->data)->cutoff_sqr;
    }
    return -1.0;
}

double fmdh_largest_cutoff(fmd_t *md)
{
    if (md->pottable == NULL)
    {
        fmdh_err_unprepared(md, false, HERE, "potential table");
        return 0.;
    }
    double m = 0.0;
    for (unsigned i = 0; i < md->nkinds; i++)
        for (unsigned j = 0; j <= i; j++)
        {
            double r = pair_cutoff_sqr(&md->pottable[i * md->nkinds + j]);
            if (r < 0)
            {
                fmdh_err_unacceptable_int(md, false, HERE, "pair potential category",
                                          md->pottable[i * md->nkinds + j].cat);
                return 0.;
            }
            if (r > m) m = r;
        }
    return sqrt(m);
}

/* _fmd_pot_update_and_process_potcats, potential.c:177-249 */
bool fmdh_update_potcats(fmd_t *md)
{
    if (md->pottable == NULL)
    {
        fmdh_err_unprepared(md, false, HERE, "potential table");
        return false;
    }
    if (md->potcats_valid) return true;
    bool has[4] = {false, false, false, false};
    int n = 0;
    for (unsigned i = 0; i < md->nkinds; i++)
        for (unsigned j = 0; j <= i; j++)
        {
            int c = md->pottable[i * md->nkinds + j].cat;
            if (c == POT_NONE)
            {
                fmdh_err_unacceptable_int(md, false, HERE, "pair potential category", c);
                return false;
            }
            if (!has[c]) { has[c] = true; n++; }
        }
    md->potcats_num = n;
    md->eam_applied = has[POT_EAM_ALLOY];
    md->potcats_valid = true;
    return true;
}

void fmdh_potsys_free(fmd_t *md)
{
    for (int i = 0; i < md->npots; i++)
    {
        if (md->potlist[i]->cat == POT_EAM_ALLOY) eam_free(md->potlist[i]->data);
        else free(md->potlist[i]->data);
        free(md->potlist[i]);
    }
    free(md->potlist);
    md->potlist = NULL;
    md->npots = 0;
    for (unsigned i = 0; i < md->nkinds; i++) free(md->kinds[i].name);
    free(md->kinds);
    md->kinds = NULL;
    free(md->pottable);
    md->pottable = NULL;
}

/* pair table and EAM tables -> device */
bool fmdh_push_potentials(fmd_t *md)
{
    unsigned n = md->nkinds;
    for (unsigned i = 0; i < n; i++)
        for (unsigned j = 0; j < n; j++)
        {
            potpair_t *p = &md->pottable[i * n + j];
            int r = FMDX_OK;
            switch (p->cat)
            {
                case POT_LJ_6_12:
                {
                    lj_t *lj = p->data;
                    r = fmdx_set_pair(md->dev, i, j, FMDX_POT_LJ, lj->cutoff_sqr, lj->eps4, lj->sig_sqr, 0.0);
                    break;
                }
                case POT_MORSE:
                {
                    morse_t *m = p->data;
                    r = fmdx_set_pair(md->dev, i, j, FMDX_POT_MORSE, m->cutoff_sqr, m->D0, m->alpha, m->r0);
                    break;
                }
                case POT_EAM_ALLOY:
                {
                    eam_t *e = p->data;
                    r = fmdx_set_pair(md->dev, i, j, FMDX_POT_EAM, e->cutoff_sqr, p->iloc, p->jloc, 0.0);
                    break;
                }
                default: break;
            }
            if (r != FMDX_OK)
            {
                fmdh_err_function_failed(md, true, HERE, "fmdx_set_pair");
                return false;
            }
        }
    if (md->eam)
    {
        eam_t *e = md->eam;
        int ne = e->nelem;
        double *F = malloc(sizeof(double) * ne * e->Nrho), *FDD = malloc(sizeof(double) * ne * e->Nrho);
        double *rho = malloc(sizeof(double) * ne * e->Nr2), *rhodd = malloc(sizeof(double) * ne * e->Nr2);
        double *phi = malloc(sizeof(double) * ne * ne * e->Nr2), *phidd = malloc(sizeof(double) * ne * ne * e->Nr2);
        for (int a = 0; a < ne; a++)
        {
            memcpy(F + (size_t)a * e->Nrho, e->el[a].F, sizeof(double) * e->Nrho);
            memcpy(FDD + (size_t)a * e->Nrho, e->el[a].Fdd, sizeof(double) * e->Nrho);
            memcpy(rho + (size_t)a * e->Nr2, e->el[a].rho, sizeof(double) * e->Nr2);
            memcpy(rhodd + (size_t)a * e->Nr2, e->el[a].rhodd, sizeof(double) * e->Nr2);
            for (int b = 0; b < ne; b++)
            {
                memcpy(phi + ((size_t)a * ne + b) * e->Nr2, e->phi[a][b], sizeof(double) * e->Nr2);
                memcpy(phidd + ((size_t)a * ne + b) * e->Nr2, e->phidd[a][b], sizeof(double) * e->Nr2);
            }
        }
        int eok[HMAXK];
        for (unsigned k = 0; k < HMAXK; k++) eok[k] = k < n ? md->kinds[k].eam_element : -1;
        int r = fmdx_set_eam(md->dev, ne, e->Nrho, e->drho, e->Nr2, e->dr2, e->cutoff_sqr, F, FDD, rho, rhodd, phi,
                             phidd, eok);
        free(F); free(FDD); free(rho); free(rhodd); free(phi); free(phidd);
        if (r != FMDX_OK)
        {
            fmdh_err_function_failed(md, true, HERE, "fmdx_set_eam");
            return false;
        }
    }
    md->dev_pot_dirty = false;
    return true;
}
