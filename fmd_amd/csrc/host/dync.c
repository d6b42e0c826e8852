/*
 * dync.c -- time loops and the host <-> device hand-over.
 *
 * The loops have the reference's structure and step count
 * (VelocityVerlet_integrate integrators.c:227-264, SymplecticEuler_integrate
 * integrators.c:181-225, fmd_dync_integrate / equilibrate integrators.c:287-325):
 * the loop runs while the accumulated FP64 time is below start + duration,
 * so the step count matches the reference for the same inputs.
 *
 * Every per-step operation is an asynchronous call into libfmdx.so.  The
 * host synchronises only when it must: a timer tick or field event whose
 * handler reads state, a getter, or the periodic check (every CHECK_EVERY
 * steps and at the end of a loop) that turns device error bits into the
 * reference's error events and grows the per-cell slot capacity before a
 * cell can outgrow it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_host.h"

#define CHECK_EVERY 32

fmd_real_t fmd_dync_getTimestep(fmd_t *md) { return md->timestep; }
fmd_real_t fmd_dync_getTime(fmd_t *md) { return md->time; }

/* ---------------------------------------------------------------- device */

/* the same rule as the device library's first choice: 1.5 x the fullest cell + 8,
   in multiples of 32 slots so that each cell's values start on a 128-byte line */
static unsigned choose_cap(unsigned maxc)
{
    unsigned c = maxc + maxc / 2 + 8;
    c = (c + 31) & ~31u;
    return c > 256 ? 256 : c;
}

static void report_device_errors(fmd_t *md, const fmdx_thermo_t *t)
{
    unsigned fresh = t->err & ~md->dev_err_reported;
    if (!fresh) return;
    md->dev_err_reported |= fresh;
    /* where the device first saw a bad position or force, for the record (the device checks every
       step; the host learns of it at the next synchronisation) */
    if (md->ShowErrorMessages && (fresh & (FMDX_DEV_ERR_POSITION | FMDX_DEV_ERR_FORCE)))
        fprintf(stderr, "FMD: device error word %u by step %d: first non-finite force in cell %d, "
                        "first unbinnable position in cell %d (-1: none)\n",
                t->err, md->time_iteration, (int)t->bad_force_cell - 1, (int)t->bad_position_cell - 1);
    if (fresh & FMDX_DEV_ERR_FORCE)
        fmdh_err_outside_real_interval(md, true, HERE, "force on an atom", NAN, "of finite values");
    if (fresh & FMDX_DEV_ERR_POSITION) fmdh_err_unexpected_position(md, true, HERE);
    if (fresh & FMDX_DEV_ERR_RHO_RANGE)
        fmdh_err_outside_real_interval(md, true, HERE, "electron density", NAN, "related to EAM potential");
    if (fresh & FMDX_DEV_ERR_OVERFLOW) fmdh_err_unable_allocate(md, true, HERE, 0);
    if (fresh & FMDX_DEV_ERR_TTM_CE)
        fmdh_err_outside_real_interval(md, true, HERE, "electron temperature", NAN, "related to electron heat capacity");
    if (fresh & FMDX_DEV_ERR_TTM_G)
        fmdh_err_outside_real_interval(md, true, HERE, "electron temperature", NAN,
                                       "related to electron-ion coupling factor");
}

static void apply_thermo(fmd_t *md, const fmdx_thermo_t *t)
{
    md->GroupPotentialEnergy = t->pe;
    md->GroupKineticEnergy = t->ke;
    md->GroupTemperature = t->temperature;
    for (int d = 0; d < 3; d++) md->GroupMomentum[d] = t->momentum[d];
    md->GroupParticlesNum = (unsigned)t->n_active;
    md->TotalNoOfParticles = t->natoms;
    md->thermo_host_valid = true;
}

void fmdh_sync_thermo(fmd_t *md)
{
    if (!md->ParticlesDistributed || md->thermo_host_valid) return;
    fmdx_thermo_t t;
    if (fmdx_read_thermo(md->dev, &t) != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_read_thermo");
        return;
    }
    apply_thermo(md, &t);
    report_device_errors(md, &t);
}

/* Periodic check of the resident state (every CHECK_EVERY steps and at the end of a loop): the
   device error word, and an audit of the cells -- the atoms counted over the cells equal the atoms
   resident, no cell holds more than its slots, every position, velocity and force is finite.  A
   broken invariant is reported once, naming the step. */
void fmdh_check_device(fmd_t *md)
{
    fmdx_thermo_t t;
    if (fmdx_audit(md->dev) != FMDX_OK || fmdx_read_thermo(md->dev, &t) != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_read_thermo");
        return;
    }
    apply_thermo(md, &t);
    report_device_errors(md, &t);
    if (!md->audit_reported && (t.audit_atoms != t.natoms || t.audit_max_cell > t.cap || t.audit_nonfinite))
    {
        char msg[200];
        snprintf(msg, sizeof msg, "device state check at step %d: %u atoms in the cells for %u resident, "
                                  "fullest cell %u of %u slots, %u atoms with non-finite x, v or F",
                 md->time_iteration, t.audit_atoms, t.natoms, t.audit_max_cell, t.cap, t.audit_nonfinite);
        md->audit_reported = true;
        fmdh_err_function_failed(md, true, HERE, msg);
    }
    if (t.max_cell_count + t.max_cell_count / 4 + 4 > t.cap && t.cap < 256)
    {
        if (fmdx_set_capacity(md->dev, choose_cap(t.max_cell_count)) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_set_capacity");
    }
}

bool fmdh_download_all(fmd_t *md, unsigned *n, double **x, double **v, int **grp, unsigned **aid, unsigned **kind,
                       unsigned **cellcnt)
{
    unsigned nn = 0, ncell = 0;
    fmdx_ncells(md->dev, &ncell);
    if (fmdx_download(md->dev, &nn, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL) != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_download");
        return false;
    }
    size_t m = nn ? nn : 1;
    *x = malloc(3 * m * sizeof(double));
    *v = malloc(3 * m * sizeof(double));
    *grp = malloc(m * sizeof(int));
    *aid = malloc(m * sizeof(unsigned));
    *kind = malloc(m * sizeof(unsigned));
    *cellcnt = malloc((ncell ? ncell : 1) * sizeof(unsigned));
    if (fmdx_download(md->dev, &nn, *cellcnt, *x, *v, NULL, NULL, *grp, *aid, *kind) != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_download");
        free(*x); free(*v); free(*grp); free(*aid); free(*kind); free(*cellcnt);
        return false;
    }
    *n = nn;
    return true;
}

/* _fmd_matt_distribute, matter.c:399-447, for one process: host grid -> HBM */
bool fmdh_distribute(fmd_t *md)
{
    if (md->ParticlesDistributed) return true;
    if (!md->grid_made && !fmdh_make_grid(md)) return false;
    if (!md->dev)
    {
        double mass[HMAXK];
        for (unsigned k = 0; k < HMAXK; k++) mass[k] = k < md->nkinds ? md->kinds[k].mass : 1.0;
        int r = fmdx_system_create(&md->dev, md->l, md->PBC, md->nc, md->cellh, md->nkinds, mass, md->precision);
        if (r != FMDX_OK)
        {
            fprintf(stderr, "FMD: cannot create the device system: %s\n", fmdx_strerror(r));
            fmdh_err_function_failed(md, true, HERE, "fmdx_system_create");
            return false;
        }
    }
    if (!fmdh_push_potentials(md)) return false;
    unsigned n = 0;
    for (unsigned ic = 0; ic < md->ngg; ic++) n += md->gg[ic].n;
    size_t m = n ? n : 1;
    double *x = malloc(3 * m * sizeof(double)), *v = malloc(3 * m * sizeof(double));
    int *grp = malloc(m * sizeof(int));
    unsigned *aid = malloc(m * sizeof(unsigned)), *kind = malloc(m * sizeof(unsigned));
    unsigned *cnt = malloc(md->ngg * sizeof(unsigned));
    unsigned k = 0;
    for (unsigned ic = 0; ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        cnt[ic] = c->n;
        memcpy(x + 3 * (size_t)k, c->x, 3 * c->n * sizeof(double));
        memcpy(v + 3 * (size_t)k, c->v, 3 * c->n * sizeof(double));
        memcpy(grp + k, c->grp, c->n * sizeof(int));
        memcpy(aid + k, c->aid, c->n * sizeof(unsigned));
        memcpy(kind + k, c->kind, c->n * sizeof(unsigned));
        k += c->n;
    }
    int r = fmdx_upload(md->dev, n, cnt, x, v, grp, aid, kind);
    free(x); free(v); free(grp); free(aid); free(kind); free(cnt);
    if (r != FMDX_OK)
    {
        fmdh_err_function_failed(md, true, HERE, "fmdx_upload");
        return false;
    }
    fmdh_gg_free(md);
    md->ParticlesDistributed = true;
    md->TotalNoOfParticles = n;
    md->KineticEnergyUpdated = false;
    md->thermo_host_valid = false;
    return true;
}

/* --------------------------------------------------------------- loops */

/* The reference sums the potential energy in every force call (forces.c:332-382); here the pair
   passes sum it only when it can be read before the next force call: when the step ends the
   loop (the caller may read it after), or when a handler may run after it -- a timer tick due at
   the step's end time, or any turi (field events) -- since a handler may call the getters. */
static bool energy_needed(fmd_t *md, double t_next, double t_end)
{
    if (!(t_next < t_end)) return true;
    if (md->EventHandler == NULL) return false;
    if (md->nturies > 0) return true;
    return fmdh_tick_due(md, t_next);
}

static void update_forces(fmd_t *md, bool energy)
{
    int fric = fmdh_ttm_friction_active(md) ? 1 : 0;
    if (fmdx_update_forces_ex(md->dev, md->ActiveGroup, fric, energy ? FMDX_FORCES_ENERGY : 0) != FMDX_OK)
        fmdh_err_function_failed(md, true, HERE, "fmdx_update_forces");
    md->thermo_host_valid = false;
}

static void ticks(fmd_t *md)
{
    if (md->EventHandler != NULL) fmdh_send_timer_ticks(md);
}

static void velocity_verlet(fmd_t *md, double duration, bool thermostat)
{
    double start = md->time;
    update_forces(md, true);
    if (md->nturies > 0) fmdh_turies_update(md, true, true, true);
    ticks(md);
    unsigned steps = 0;
    while (md->time < start + duration)
    {
        if (thermostat && !md->KineticEnergyUpdated)
        {
            fmdx_thermo(md->dev, md->ActiveGroup);   /* integrators.c:46 */
            md->KineticEnergyUpdated = true;
        }
        if (fmdx_vv_start(md->dev, md->ActiveGroup, md->timestep, thermostat, md->BerendsenThermostatParam,
                          md->DesiredTemperature) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_vv_start");
        update_forces(md, energy_needed(md, md->time + md->timestep, start + duration));
        if (fmdx_vv_finish(md->dev, md->ActiveGroup, md->timestep) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_vv_finish");
        md->KineticEnergyUpdated = true;
        md->thermo_host_valid = false;
        md->time += md->timestep;
        md->time_iteration++;
        if (md->nturies > 0) fmdh_turies_update(md, true, true, true);
        ticks(md);
        if (++steps % CHECK_EVERY == 0) fmdh_check_device(md);
    }
    fmdh_check_device(md);
}

static void symplectic_euler(fmd_t *md, double duration)
{
    double start = md->time;
    if (md->active_ttm_turi >= 0) fmdh_ttm_get_ready(md);
    if (md->nturies > 0) fmdh_turies_update(md, true, true, false);
    update_forces(md, true);
    if (md->nturies > 0) fmdh_turies_update(md, false, false, true);
    ticks(md);
    unsigned steps = 0;
    while (md->time < start + duration)
    {
        if (fmdx_se_step(md->dev, md->ActiveGroup, md->timestep) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_se_step");
        md->KineticEnergyUpdated = true;
        md->thermo_host_valid = false;
        md->time += md->timestep;
        md->time_iteration++;
        if (md->nturies > 0) fmdh_turies_update(md, true, true, false);
        update_forces(md, energy_needed(md, md->time, start + duration));
        if (md->nturies > 0) fmdh_turies_update(md, false, false, true);
        ticks(md);
        if (++steps % CHECK_EVERY == 0) fmdh_check_device(md);
    }
    fmdh_check_device(md);
}

static void set_active_group(fmd_t *md, int g)
{
    if (g != md->ActiveGroup)
    {
        md->ActiveGroup = g;
        md->KineticEnergyUpdated = false;
        md->thermo_host_valid = false;
    }
}

static bool prepare(fmd_t *md)
{
    if (!fmdh_update_potcats(md)) return false;
    if (!md->ParticlesDistributed && !fmdh_distribute(md)) return false;
    if (md->dev_pot_dirty && !fmdh_push_potentials(md)) return false;
    return true;
}

void fmd_dync_integrate(fmd_t *md, int GroupID, fmd_real_t duration, fmd_real_t timestep)
{
    if (!prepare(md)) return;
    set_active_group(md, GroupID);
    md->timestep = timestep;
    if (md->active_ttm_turi >= 0)
        symplectic_euler(md, duration);
    else
        velocity_verlet(md, duration, false);
}

void fmd_dync_equilibrate(fmd_t *md, int GroupID, fmd_real_t duration, fmd_real_t timestep, fmd_real_t tau,
                          fmd_real_t temperature)
{
    if (!prepare(md)) return;
    double bak_time = md->time;
    md->time = 0.0;
    md->timestep = timestep;
    md->DesiredTemperature = temperature;
    md->BerendsenThermostatParam = tau;
    set_active_group(md, GroupID);
    velocity_verlet(md, duration, true);
    md->time = bak_time;
}
