/*
 * turi.c -- turies, their fields and the 1-D two-temperature model.
 *
 * Reference behaviour followed:
 *   fmd_turi_add geometry          turi.c:648-738 (one process: every turi
 *                                  cell owned here, tcellh = L / tdims)
 *   field registration / deps      turi.c:1382-1500, ttm.c:385-429
 *                                  (TTM turi fields in index order: NUMBER,
 *                                  MASS, VCM, TEMPERATURE, TTM_TE, TTM_XI)
 *   update order and events        turi.c:1509-1665, 959-1373 (fields are
 *                                  visited from the highest index down; a
 *                                  field pulls its dependencies first; each
 *                                  update raises FMD_EVENT_FIELD_UPDATE)
 *   TTM setters, laser             ttm.c:549-816
 *   getReady                       ttm.c:63-80
 *   getArray                       turi.c:1675-1721
 * The sums over atoms, the presolve and the FD sub-steps run on the device
 * (fmdx_ttm_*); the host walks the field list to raise the events in the
 * reference's order, downloading arrays only when a handler asks for them.
 * TTM turies are computed: 1 x 1 x n cells with the reference's 1-D
 * update, and type 1 grids with the 3-D update of the electron temperature
 * (the reference's own 3-D updaters are absent, ttm.c:449-453, 489-493).
 * Other (custom) turi fields report FMD_ERR_NOT_SUPPORTED_YET.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_host.h"

static bool time_within(fmd_t *md, const turi_t_ *t)
{
    /* _is_time_within_turi_start_stop_times, turi.c:38-42 */
    return (md->time >= t->starttime - md->timestep / 2.0) &&
           ((md->time < t->stoptime + md->timestep / 2.0) || t->stoptime < t->starttime);
}

static int field_add(turi_t_ *t, int cat, double interval, bool allhave);

static void intervals_add(field_t_ *f, double interval, bool allhave)
{
    for (int j = 0; j < f->nint_all; j++)
        if (f->intervals_all[j] == interval) return;
    for (int j = 0; j < f->nint; j++)
        if (f->intervals[j] == interval)
        {
            if (allhave && f->nint_all < 16)
            {
                f->intervals[j] = f->intervals[--f->nint];
                f->intervals_all[f->nint_all++] = interval;
            }
            return;
        }
    if (allhave) { if (f->nint_all < 16) f->intervals_all[f->nint_all++] = interval; }
    else if (f->nint < 16) f->intervals[f->nint++] = interval;
}

/* _fmd_field_add, turi.c:1420-1502 */
static int field_add(turi_t_ *t, int cat, double interval, bool allhave)
{
    int d1 = -1, d2 = -1;
    switch (cat)
    {
        case FMD_FIELD_TEMPERATURE:
            d1 = field_add(t, FMD_FIELD_NUMBER, interval, allhave);
            d2 = field_add(t, FMD_FIELD_VCM, interval, true);
            break;
        case FMD_FIELD_VCM: d1 = field_add(t, FMD_FIELD_MASS, interval, allhave); break;
        case FMD_FIELD_NUMBER_DENSITY: d1 = field_add(t, FMD_FIELD_NUMBER, interval, allhave); break;
        case FMD_FIELD_TTM_TE:
        case FMD_FIELD_TTM_XI: d1 = field_add(t, FMD_FIELD_TEMPERATURE, interval, false); break;
        default: break;
    }
    int i;
    for (i = 0; i < t->nfields; i++)
        if (t->fields[i].cat == cat) break;
    if (i == t->nfields)
    {
        if (t->nfields >= 16) return -1;
        field_t_ *f = &t->fields[t->nfields++];
        memset(f, 0, sizeof *f);
        f->cat = cat;
        f->timestep = -1;
        f->dep[0] = d1;
        f->dep[1] = d2;
        f->ndep = cat == FMD_FIELD_TEMPERATURE ? 2 : (d1 >= 0 ? 1 : 0);
    }
    intervals_add(&t->fields[i], interval, allhave);
    return i;
}

fmd_handle_t fmd_turi_add(fmd_t *md, fmd_turi_t cat, int dimx, int dimy, int dimz, fmd_real_t starttime,
                          fmd_real_t stoptime)
{
    if (cat != FMD_TURI_TTM_TYPE1 && cat != FMD_TURI_TTM_TYPE2)
    {
        fmdh_err_not_supported(md, false, HERE, "Only TTM turies are computed by this build.");
        return -1;
    }
    if (dimx < 1 || dimy < 1 || dimz < 1)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "turi dimension", dimx < 1 ? dimx : (dimy < 1 ? dimy : dimz));
        return -1;
    }
    if (dimx == 1 && dimy == 1 ? dimz > FMDX_TTM_MAX_CELLS
                               : (cat != FMD_TURI_TTM_TYPE1 ||
                                  (long)dimx * dimy * (dimz + 2) > FMDX_TTM_MAX_TCELLS))
    {
        fmdh_err_not_supported(md, false, HERE,
                               "TTM turies: 1 x 1 x n (n <= 1024), or type 1 with dimx*dimy*(dimz+2) <= 4000.");
        return -1;
    }
    turi_t_ *nt = realloc(md->turies, (md->nturies + 1) * sizeof(turi_t_));
    if (!nt) return -1;
    md->turies = nt;
    turi_t_ *t = &md->turies[md->nturies];
    memset(t, 0, sizeof *t);
    t->cat = cat;
    t->tdims[0] = dimx; t->tdims[1] = dimy; t->tdims[2] = dimz;
    t->starttime = starttime;
    t->stoptime = stoptime;
    for (int d = 0; d < 3; d++) t->tcellh[d] = md->l[d] / t->tdims[d];
    t->tcell_volume = t->tcellh[0] * t->tcellh[1] * t->tcellh[2];
    t->is_ttm = true;
    t->ttm_type = cat == FMD_TURI_TTM_TYPE1 ? 1 : 2;
    /* init_common_types_1_2, ttm.c:385-429 */
    t->inum = field_add(t, FMD_FIELD_NUMBER, md->timestep, false);
    t->ivcm = field_add(t, FMD_FIELD_VCM, md->timestep, true);
    t->iTi = field_add(t, FMD_FIELD_TEMPERATURE, md->timestep, false);
    t->iTe = field_add(t, FMD_FIELD_TTM_TE, md->timestep, false);
    t->ixi = field_add(t, FMD_FIELD_TTM_XI, md->timestep, true);
    t->imass = t->fields[t->ivcm].dep[0];
    t->CellActivFrac = 0.1;
    t->ratio = 1;
    t->ncells = (unsigned)(dimx * dimy * (dimz + 2));
    t->Te_host = calloc(t->ncells, sizeof(double));
    md->active_ttm_turi = (int)md->nturies;
    return (fmd_handle_t)md->nturies++;
}

fmd_handle_t fmd_field_find(fmd_t *md, fmd_handle_t turi, fmd_field_t cat)
{
    if (turi < 0 || (unsigned)turi >= md->nturies) return -1;
    turi_t_ *t = &md->turies[turi];
    for (int i = 0; i < t->nfields; i++)
        if (t->fields[i].cat == (int)cat) return i;
    return -1;
}

fmd_handle_t fmd_field_add(fmd_t *md, fmd_handle_t turi, fmd_field_t cat, fmd_real_t interval)
{
    if (turi < 0 || (unsigned)turi >= md->nturies) return -1;
    return field_add(&md->turies[turi], cat, interval, false);
}

void fmd_field_save_as_hdf5(fmd_t *md, fmd_handle_t turi, fmd_handle_t field, fmd_string_t filename)
{
    (void)turi; (void)field; (void)filename;
    fmdh_err_not_supported(md, false, HERE, "HDF5 output is not part of this build.");
}

static turi_t_ *ttm_of(fmd_t *md, fmd_handle_t turi, bool type1, bool type2)
{
    if (turi < 0 || (unsigned)turi >= md->nturies) return NULL;
    turi_t_ *t = &md->turies[turi];
    if ((t->ttm_type == 1 && !type1) || (t->ttm_type == 2 && !type2))
    {
        fmdh_err_unacceptable_int(md, false, HERE, "turi category", t->cat);
        return NULL;
    }
    return t;
}

/* -------------------------------------------------------------- setters */

static bool read_table(fmd_t *md, const char *path, double tscale, double vscale, double **T, double **V,
                       double **DD, int *n)
{
    FILE *fp = fopen(path, "r");
    if (!fp)
    {
        fmdh_err_unable_open_file(md, false, HERE, path);
        return false;
    }
    int cap = 0, i = 0;
    double a, b;
    double *tt = NULL, *vv = NULL;
    while (fscanf(fp, "%lf%lf", &a, &b) == 2)
    {
        if (i == cap)
        {
            cap += 100;
            tt = realloc(tt, cap * sizeof(double));
            vv = realloc(vv, cap * sizeof(double));
        }
        tt[i] = a * tscale;
        vv[i] = b * vscale;
        i++;
    }
    fclose(fp);
    free(*T); free(*V); free(*DD);
    *T = tt;
    *V = vv;
    *n = i;
    *DD = malloc((i ? i : 1) * sizeof(double));
    if (i >= 2) fmdh_spline_prepare_adv(tt, vv, i, *DD);
    return true;
}

/* ttm.c:549-606 */
void _fmd_ttm_setHeatCapacity_file(fmd_t *md, fmd_handle_t turi, fmd_string_t path)
{
    turi_t_ *t = ttm_of(md, turi, false, true);
    if (t) read_table(md, path, 1e4, 1e5 * JOULE_PER_METER3_KELVIN, &t->T_C, &t->Ct, &t->C_DD, &t->nC);
}

void _fmd_ttm_setHeatCapacity_linear(fmd_t *md, fmd_handle_t turi, fmd_ttm_heat_capacity_linear_t c)
{
    turi_t_ *t = ttm_of(md, turi, true, false);
    if (t) t->C_gamma = c.gamma * JOULE_PER_METER3_KELVIN2;
}

void _fmd_ttm_setHeatConductivity_zhigilei(fmd_t *md, fmd_handle_t turi, fmd_ttm_heat_conductivity_zhigilei_t k)
{
    turi_t_ *t = ttm_of(md, turi, false, true);
    if (!t) return;
    t->Kzh_A = k.A / SECOND;
    t->Kzh_B = k.B / SECOND;
    t->Kzh_v2 = (k.v * METER_PER_SECOND) * (k.v * METER_PER_SECOND);
}

void _fmd_ttm_setHeatConductivity_constant1(fmd_t *md, fmd_handle_t turi, fmd_real_t k)
{
    turi_t_ *t = ttm_of(md, turi, true, false);
    if (t)
    {
        t->K = k * WATT_PER_METER_KELVIN;
        t->kmode = 0;
    }
}

void fmd_ext_ttm_setHeatConductivity_proportional(fmd_t *md, fmd_handle_t turi, fmd_real_t K0)
{
    turi_t_ *t = ttm_of(md, turi, true, false);
    if (!t) return;
    t->K = K0 * WATT_PER_METER_KELVIN;
    t->kmode = 1;
}

void fmd_ext_ttm_useGridSolver(fmd_t *md, fmd_handle_t turi, bool on)
{
    turi_t_ *t = ttm_of(md, turi, true, false);
    if (t) t->grid_solver = on;
}

void _fmd_ttm_setHeatConductivity_constant2(fmd_t *md, fmd_handle_t turi, fmd_ttm_heat_conductivity_constant_t k)
{
    _fmd_ttm_setHeatConductivity_constant1(md, turi, k.value);
}

void _fmd_ttm_setCouplingFactor_file(fmd_t *md, fmd_handle_t turi, fmd_string_t path)
{
    turi_t_ *t = ttm_of(md, turi, false, true);
    if (t) read_table(md, path, 1e4, 1e17 * WATT_PER_METER3_KELVIN, &t->T_G, &t->Gt, &t->G_DD, &t->nG);
}

void _fmd_ttm_setCouplingFactor_constant1(fmd_t *md, fmd_handle_t turi, fmd_real_t g)
{
    turi_t_ *t = ttm_of(md, turi, true, false);
    if (t) t->G = g * WATT_PER_METER3_KELVIN;
}

void _fmd_ttm_setCouplingFactor_constant2(fmd_t *md, fmd_handle_t turi, fmd_ttm_coupling_factor_constant_t g)
{
    _fmd_ttm_setCouplingFactor_constant1(md, turi, g.value);
}

void fmd_ttm_setElectronTemperature(fmd_t *md, fmd_handle_t turi, fmd_real_t Te)
{
    turi_t_ *t = ttm_of(md, turi, true, true);
    if (!t) return;
    const unsigned nzg = t->tdims[2] + 2;
    for (unsigned a = 0; a < t->ncells; a++)
        if (a % nzg >= 1 && a % nzg <= t->tdims[2]) t->Te_host[a] = Te;
    if (t->dev_ready && md->dev)
    {
        double *buf = malloc(t->ncells * sizeof(double));
        fmdx_ttm_download(md->dev, buf, NULL, NULL, NULL, NULL, NULL);
        for (unsigned a = 0; a < t->ncells; a++)
            if (a % nzg >= 1 && a % nzg <= t->tdims[2]) buf[a] = Te;
        fmdx_ttm_upload_te(md->dev, buf);
        free(buf);
    }
}

void fmd_ttm_setTimestepRatio(fmd_t *md, fmd_handle_t turi, int ratio)
{
    turi_t_ *t = ttm_of(md, turi, true, true);
    if (t) t->ratio = ratio;
}

void fmd_ttm_setCellActivationFraction(fmd_t *md, fmd_handle_t turi, fmd_real_t value)
{
    turi_t_ *t = ttm_of(md, turi, true, true);
    if (!t) return;
    if (value < 0 || value > 1)
    {
        fmdh_err_outside_real_interval(md, false, HERE, "value", value, "[0 1]");
        return;
    }
    t->CellActivFrac = value;
}

/* ttm.c:794-816 */
void _fmd_ttm_setLaserSource_gaussian(fmd_t *md, fmd_handle_t turi, fmd_ttm_laser_gaussian_t laser)
{
    turi_t_ *t = ttm_of(md, turi, true, true);
    if (!t) return;
    double Lp = laser.AbsorptionDepth * METER;
    double R = laser.reflectance;
    double I0 = sqrt(4 * log(2) / M_PI) * laser.fluence / laser.duration * WATT_PER_METER2;
    double sigma = laser.duration / sqrt(8 * log(2)) * SECOND;
    t->laser_t0 = laser.t0 * SECOND;
    t->laser_tdiff = 4 * sqrt(2) * sigma;
    t->laser_2sig2_inv = -0.5 / (sigma * sigma);
    t->laser_absdepth_inv = -1. / Lp;
    t->laser_factor = I0 * (1 - R) / Lp;
}

/* --------------------------------------------------------------- device */

static bool ttm_device_setup(fmd_t *md, turi_t_ *t)
{
    fmdx_ttm_t c;
    memset(&c, 0, sizeof c);
    c.type = t->ttm_type;
    c.ncz = (int)t->tdims[2];
    c.ratio = t->ratio;
    c.tcellh_z = t->tcellh[2];
    c.tcell_volume = t->tcell_volume;
    c.C_gamma = t->C_gamma; c.K = t->K; c.G = t->G;
    c.Kzh_A = t->Kzh_A; c.Kzh_B = t->Kzh_B; c.Kzh_v2 = t->Kzh_v2;
    c.nC = t->nC; c.nG = t->nG;
    c.T_C = t->T_C; c.Ct = t->Ct; c.C_DD = t->C_DD;
    c.T_G = t->T_G; c.Gt = t->Gt; c.G_DD = t->G_DD;
    c.laser_factor = t->laser_factor; c.laser_absdepth_inv = t->laser_absdepth_inv;
    c.laser_2sig2_inv = t->laser_2sig2_inv; c.laser_t0 = t->laser_t0; c.laser_tdiff = t->laser_tdiff;
    c.frontsurf = t->frontsurf;
    c.min_atoms = t->min_atoms;
    c.ncx = (int)t->tdims[0];
    c.ncy = (int)t->tdims[1];
    c.tcellh_x = t->tcellh[0];
    c.tcellh_y = t->tcellh[1];
    c.kmode = t->kmode;
    c.solver3d = t->grid_solver;
    if (fmdx_ttm_setup(md->dev, &c, t->Te_host) != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_ttm_setup");
        return false;
    }
    t->dev_ready = true;
    return true;
}

/* _fmd_ttm_getReady, ttm.c:63-80 */
void fmdh_ttm_get_ready(fmd_t *md)
{
    if (md->active_ttm_turi < 0) return;
    turi_t_ *t = &md->turies[md->active_ttm_turi];
    fmd_rtriple_t ll, ul;
    fmd_matt_findLimits(md, md->ActiveGroup, ll, ul);
    t->frontsurf = ll[2];
    double vol = 1.0;
    for (int d = 0; d < 3; d++) vol *= ul[d] - ll[d];
    (void)fmd_matt_getTemperature(md);   /* makes GroupParticlesNum current */
    t->initial_atoms = (unsigned)(md->GroupParticlesNum / vol * t->tcell_volume);
    t->min_atoms = (unsigned)(t->initial_atoms * t->CellActivFrac);
    if (!t->dev_ready) ttm_device_setup(md, t);
    else fmdx_ttm_set_frontsurf(md->dev, t->frontsurf, t->min_atoms);
}

bool fmdh_ttm_friction_active(fmd_t *md)
{
    if (md->active_ttm_turi < 0) return false;
    turi_t_ *t = &md->turies[md->active_ttm_turi];
    return t->dev_ready && time_within(md, t);
}

/* ----------------------------------------------------- update and events */

static void field_event(fmd_t *md, int turi, int field)
{
    if (md->EventHandler == NULL) return;
    fmd_event_params_field_update_t p;
    p.turi = turi;
    p.field = field;
    md->EventHandler(md, FMD_EVENT_FIELD_UPDATE, md->userobject, (fmd_params_t *)&p);
}

static void mark(fmd_t *md, field_t_ *f)
{
    f->timestep = md->time_iteration;
    f->time = md->time;
}

typedef struct
{
    fmd_t *md;
    turi_t_ *t;
    int ti;
    bool sums_done;
} walk_t;

static void ensure_sums(walk_t *w)
{
    if (w->sums_done) return;
    if (!w->t->dev_ready) ttm_device_setup(w->md, w->t);
    if (fmdx_ttm_fields(w->md->dev) != FMDX_OK) fmdh_err_function_failed(w->md, true, HERE, "fmdx_ttm_fields");
    w->sums_done = true;
}

static bool updated(walk_t *w, int fi) { return w->t->fields[fi].timestep == w->md->time_iteration; }

static void upd_number(walk_t *w, int fi)
{
    ensure_sums(w);
    mark(w->md, &w->t->fields[fi]);
    field_event(w->md, w->ti, fi);
}

/* update_field_vcm (turi.c:1180-1188) */
static void upd_vcm(walk_t *w, int fi)
{
    ensure_sums(w);
    int fm = w->t->fields[fi].dep[0];
    if (!updated(w, fm))
    {
        mark(w->md, &w->t->fields[fm]);
        field_event(w->md, w->ti, fm);
    }
    mark(w->md, &w->t->fields[fi]);
    field_event(w->md, w->ti, fi);
}

/* update_field_temperature (turi.c:1322-1335) */
static void upd_temperature(walk_t *w, int fi)
{
    field_t_ *f = &w->t->fields[fi];
    int fnum = f->dep[0], fvcm = f->dep[1];
    if (!updated(w, fvcm)) upd_vcm(w, fvcm);
    ensure_sums(w);
    if (!updated(w, fnum))
    {
        mark(w->md, &w->t->fields[fnum]);
        field_event(w->md, w->ti, fnum);
    }
    mark(w->md, f);
    field_event(w->md, w->ti, fi);
}

/* update_field_ttm_Te_and_xi (turi.c:1337-1373) */
static void upd_ttm(walk_t *w)
{
    fmd_t *md = w->md;
    turi_t_ *t = w->t;
    if (!updated(w, t->iTi)) upd_temperature(w, t->iTi);
    if (fmdx_ttm_presolve(md->dev) != FMDX_OK) fmdh_err_function_failed(md, true, HERE, "fmdx_ttm_presolve");
    t->fields[t->iTe].time = md->time;
    field_event(md, w->ti, t->iTe);
    if (fmdx_ttm_solve(md->dev, md->time, md->timestep) != FMDX_OK)
        fmdh_err_function_failed(md, true, HERE, "fmdx_ttm_solve");
    mark(md, &t->fields[t->ixi]);
    t->fields[t->iTe].time = md->time + md->timestep;
    t->fields[t->iTe].timestep = md->time_iteration;
}

/* _fmd_turies_update, turi.c:1627-1665 */
void fmdh_turies_update(fmd_t *md, bool X, bool V, bool F)
{
    for (unsigned ti = 0; ti < md->nturies; ti++)
    {
        turi_t_ *t = &md->turies[ti];
        if (!time_within(md, t)) continue;
        walk_t w = {md, t, (int)ti, false};
        for (int fi = t->nfields - 1; fi >= 0; fi--)
        {
            field_t_ *f = &t->fields[fi];
            if (f->timestep == md->time_iteration) continue;
            bool all = X && V && F;
            if (!all && !(X && V))
            {
                fmdh_err_not_supported(md, false, HERE, "This combination of input values is not supported yet!");
                continue;
            }
            switch (f->cat)
            {
                case FMD_FIELD_TTM_TE:
                case FMD_FIELD_TTM_XI:
                    if (!all) upd_ttm(&w);
                    break;
                case FMD_FIELD_TEMPERATURE: upd_temperature(&w, fi); break;
                case FMD_FIELD_VCM: upd_vcm(&w, fi); break;
                case FMD_FIELD_NUMBER_DENSITY:
                    if (!updated(&w, f->dep[0])) upd_number(&w, f->dep[0]);
                    mark(md, f);
                    field_event(md, (int)ti, fi);
                    break;
                case FMD_FIELD_NUMBER: upd_number(&w, fi); break;
                case FMD_FIELD_MASS: upd_number(&w, fi); break;
            }
        }
    }
}

/* fmd_field_getArray, turi.c:1675-1721 */
fmd_array3s_t *fmd_field_getArray(fmd_t *md, fmd_handle_t turi, fmd_handle_t field, fmd_array3_t *array,
                                  fmd_utriple_t dims)
{
    if (turi < 0 || (unsigned)turi >= md->nturies) return NULL;
    turi_t_ *t = &md->turies[turi];
    if (field < 0 || field >= t->nfields) return NULL;
    int cat = t->fields[field].cat;
    const unsigned nx = t->tdims[0], ny = t->tdims[1], nz = t->tdims[2], nzg = nz + 2, n = t->ncells;
    if (!t->dev_ready || !md->dev)
    {
        fmdh_err_unprepared(md, false, HERE, "turi field");
        return NULL;
    }
    double *Te = malloc(n * sizeof(double)), *Ti = malloc(n * sizeof(double));
    double *vcm = malloc(3 * n * sizeof(double)), *mass = malloc(n * sizeof(double));
    unsigned *num = malloc(n * sizeof(unsigned));
    fmdx_ttm_download(md->dev, Te, Ti, NULL, vcm, num, mass);
    fmd_array3s_t *a = NULL;
    size_t esz = cat == FMD_FIELD_VCM ? 3 * sizeof(double) : (cat == FMD_FIELD_NUMBER ? sizeof(unsigned)
                                                                                      : sizeof(double));
    switch (cat)
    {
        case FMD_FIELD_VCM: case FMD_FIELD_NUMBER_DENSITY: case FMD_FIELD_TEMPERATURE:
        case FMD_FIELD_TTM_TE: case FMD_FIELD_MASS: case FMD_FIELD_NUMBER:
            a = fmdh_array3s_create(nx, ny, nz, esz);
            break;
        default:
            fmdh_err_unacceptable_int(md, false, HERE, "field category", cat);
    }
    if (a)
        for (unsigned ix = 0; ix < nx; ix++)
            for (unsigned iy = 0; iy < ny; iy++)
                for (unsigned iz = 0; iz < nz; iz++)
                {
                    const unsigned s_ = (ix * ny + iy) * nzg + iz + 1;
                    char *row = (char *)((void ***)a->data)[ix][iy];
                    switch (cat)
                    {
                        case FMD_FIELD_VCM:
                            for (int d = 0; d < 3; d++) ((double *)row)[3 * iz + d] = vcm[3 * s_ + d];
                            break;
                        case FMD_FIELD_NUMBER_DENSITY: ((double *)row)[iz] = num[s_] / t->tcell_volume; break;
                        case FMD_FIELD_TEMPERATURE: ((double *)row)[iz] = Ti[s_]; break;
                        case FMD_FIELD_TTM_TE: ((double *)row)[iz] = Te[s_]; break;
                        case FMD_FIELD_MASS: ((double *)row)[iz] = mass[s_] * MD_MASS_UNIT; break;
                        case FMD_FIELD_NUMBER: ((unsigned *)row)[iz] = num[s_]; break;
                    }
                }
    free(Te); free(Ti); free(vcm); free(mass); free(num);
    if (a)
    {
        *array = a->data;
        dims[0] = a->dims[0];
        dims[1] = a->dims[1];
        dims[2] = a->dims[2];
    }
    return a;
}

void fmdh_turi_free_all(fmd_t *md)
{
    for (unsigned i = 0; i < md->nturies; i++)
    {
        turi_t_ *t = &md->turies[i];
        free(t->T_C); free(t->Ct); free(t->C_DD); free(t->T_G); free(t->Gt); free(t->G_DD);
        free(t->Te_host);
    }
    free(md->turies);
    md->turies = NULL;
    md->nturies = 0;
    md->active_ttm_turi = -1;
}
