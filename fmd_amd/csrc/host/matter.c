/*
 * matter.c -- the global cell grid before distribution, lattice builders,
 * velocity edits, group quantities and configuration output.
 *
 * Reference behaviour followed:
 *   grid geometry          misc.c:41-62 (nc = (int)(L / rc_max), cellh = L / nc)
 *   cell append / remove   cell.h:84-95, cell.c:151-160
 *   cuboid builders        structure.c:35-265 (quarter-cell offset, random
 *                          kinds by ratio, Gaussian velocities, remaining
 *                          momentum removed per group)
 *   Maxwell velocities     matter.c:588-633
 *   translate              matter.c:482-552
 *   addVelocity            matter.c:117-150
 *   findLimits             matter.c:152-213
 *   changeGroupID          matter.c:449-480
 *   group thermodynamics   matter.c:31-115, 568-648
 *   saveConfiguration      matter.c:650-829
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_host.h"

#define RESERVED_GROUP -2

/* ------------------------------------------------------------ host cells */

bool fmdh_make_grid(fmd_t *md)
{
    if (md->grid_made) return true;
    double rc = fmdh_largest_cutoff(md);
    if (!(rc > 0)) return false;
    for (int d = 0; d < 3; d++)
    {
        md->nc[d] = (int)(md->l[d] / rc);
        if (md->nc[d] < 1)
        {
            fmdh_err_unacceptable_int(md, true, HERE, "number of grid cells", md->nc[d]);
            return false;
        }
        md->cellh[d] = md->l[d] / md->nc[d];
    }
    md->ngg = (unsigned)(md->nc[0] * md->nc[1] * md->nc[2]);
    md->gg = calloc(md->ngg, sizeof(hcell_t));
    if (!md->gg)
    {
        fmdh_err_unable_allocate(md, true, HERE, md->ngg * sizeof(hcell_t));
        return false;
    }
    md->grid_made = true;
    return true;
}

unsigned fmdh_cell_new(fmd_t *md, hcell_t *c)
{
    (void)md;
    if (c->n == c->cap)
    {
        unsigned cap = c->cap ? 2 * c->cap : 8;
        c->x = realloc(c->x, 3 * cap * sizeof(double));
        c->v = realloc(c->v, 3 * cap * sizeof(double));
        c->grp = realloc(c->grp, cap * sizeof(int));
        c->aid = realloc(c->aid, cap * sizeof(unsigned));
        c->kind = realloc(c->kind, cap * sizeof(unsigned));
        c->cap = cap;
    }
    return c->n++;
}

static void cell_copy(hcell_t *from, unsigned i, hcell_t *to, unsigned j)
{
    for (int d = 0; d < 3; d++)
    {
        to->x[3 * j + d] = from->x[3 * i + d];
        to->v[3 * j + d] = from->v[3 * i + d];
    }
    to->grp[j] = from->grp[i];
    to->aid[j] = from->aid[i];
    to->kind[j] = from->kind[i];
}

void fmdh_cell_remove(hcell_t *c, unsigned i)
{
    c->n--;
    if (c->n != i) cell_copy(c, c->n, c, i);
}

void fmdh_gg_free(fmd_t *md)
{
    if (!md->gg) return;
    for (unsigned i = 0; i < md->ngg; i++)
    {
        hcell_t *c = &md->gg[i];
        free(c->x); free(c->v); free(c->grp); free(c->aid); free(c->kind);
    }
    free(md->gg);
    md->gg = NULL;
}

static inline unsigned flat3(const int *n, const int *i) { return (unsigned)(i[2] + n[2] * (i[1] + n[1] * i[0])); }

/* -------------------------------------------------------------- builders */

typedef enum { LATTICE_FCC, LATTICE_BCC, LATTICE_SC } lattice_t;

static void make_cuboid(fmd_t *md, lattice_t lt, double x, double y, double z, int dimx, int dimy, int dimz,
                        double lp, const double *ratio, int GroupID, double temp)
{
    static const double r_fcc[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};
    static const double r_bcc[2][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}};
    static const double r_sc[1][3] = {{0.0, 0.0, 0.0}};
    const double (*rp)[3];
    int ppc;
    switch (lt)
    {
        case LATTICE_FCC: rp = r_fcc; ppc = 4; break;
        case LATTICE_BCC: rp = r_bcc; ppc = 2; break;
        default: rp = r_sc; ppc = 1; break;
    }
    double cum[HMAXK], sum = 0.0;
    for (unsigned i = 0; i < md->nkinds; i++) cum[i] = (sum += ratio[i]);
    double P[3] = {0.0, 0.0, 0.0};
    int dims[3] = {dimx, dimy, dimz};
    double r0[3] = {x, y, z};
    unsigned added = 0;
    int cc[3];
    for (cc[0] = 0; cc[0] < dims[0]; cc[0]++)
        for (cc[1] = 0; cc[1] < dims[1]; cc[1]++)
            for (cc[2] = 0; cc[2] < dims[2]; cc[2]++)
                for (int i = 0; i < ppc; i++)
                {
                    double px[3];
                    int ic[3];
                    bool inside = true;
                    for (int d = 0; d < 3; d++)
                    {
                        px[d] = r0[d] + (cc[d] + .25 + rp[i][d]) * lp;
                        ic[d] = (int)floor(px[d] / md->cellh[d]);
                        if (ic[d] < 0 || ic[d] >= md->nc[d]) inside = false;
                    }
                    double rn = sum * fmdh_rng_uniform(md);
                    unsigned j;
                    for (j = 0; j < md->nkinds; j++)
                        if (rn < cum[j]) break;
                    if (j == md->nkinds) j = md->nkinds - 1;
                    double mass = md->kinds[j].mass;
                    double sd = sqrt(K_BOLTZMANN * temp / mass);
                    double vel[3];
                    for (int d = 0; d < 3; d++) vel[d] = fmdh_rng_gaussian(md, sd);
                    if (!inside)
                    {
                        fmdh_err_unexpected_position(md, false, HERE);
                        continue;
                    }
                    hcell_t *c = &md->gg[flat3(md->nc, ic)];
                    unsigned pi = fmdh_cell_new(md, c);
                    c->kind[pi] = j;
                    c->grp[pi] = RESERVED_GROUP;
                    c->aid[pi] = md->TotalNoOfParticles++;
                    for (int d = 0; d < 3; d++)
                    {
                        c->x[3 * pi + d] = px[d];
                        c->v[3 * pi + d] = vel[d];
                        P[d] += mass * vel[d];
                    }
                    added++;
                }
    /* removeRemainingMomentum, structure.c:32-50 */
    unsigned natoms = (unsigned)ppc * dimx * dimy * dimz;
    for (unsigned ic = 0; ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        for (unsigned i = 0; i < c->n; i++)
            if (c->grp[i] == RESERVED_GROUP)
            {
                c->grp[i] = GroupID;
                double mass = md->kinds[c->kind[i]].mass;
                for (int d = 0; d < 3; d++) c->v[3 * i + d] -= P[d] / (natoms * mass);
            }
    }
    (void)added;
}

static bool cuboid_prologue(fmd_t *md, int GroupID)
{
    if (md->ParticlesDistributed)
    {
        fmdh_err_not_supported(md, false, HERE, "Matter cannot be added after the particles are distributed.");
        return false;
    }
    if (!md->grid_made && !fmdh_make_grid(md)) return false;
    if (GroupID == md->ActiveGroup || md->ActiveGroup == FMD_GROUP_ALL) md->KineticEnergyUpdated = false;
    if (GroupID < 0)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "GroupID", GroupID);
        return false;
    }
    return true;
}

#define CUBOID_SINGLE(NAME, LT)                                                                          \
    void NAME(fmd_t *md, fmd_real_t x, fmd_real_t y, fmd_real_t z, int dimx, int dimy, int dimz,        \
              fmd_real_t lp, unsigned atomkind, int GroupID, fmd_real_t temp)                            \
    {                                                                                                    \
        if (!cuboid_prologue(md, GroupID)) return;                                                       \
        double ratio[HMAXK] = {0};                                                                       \
        if (atomkind < md->nkinds) ratio[atomkind] = 1.0;                                                \
        make_cuboid(md, LT, x, y, z, dimx, dimy, dimz, lp, ratio, GroupID, temp);                        \
    }

#define CUBOID_MIX(NAME, LT)                                                                             \
    void NAME(fmd_t *md, fmd_real_t x, fmd_real_t y, fmd_real_t z, int dimx, int dimy, int dimz,        \
              fmd_real_t lp, fmd_real_t ratio[], int GroupID, fmd_real_t temp)                           \
    {                                                                                                    \
        if (!cuboid_prologue(md, GroupID)) return;                                                       \
        make_cuboid(md, LT, x, y, z, dimx, dimy, dimz, lp, ratio, GroupID, temp);                        \
    }

CUBOID_SINGLE(fmd_matt_makeCuboidSC, LATTICE_SC)
CUBOID_SINGLE(fmd_matt_makeCuboidBCC, LATTICE_BCC)
CUBOID_SINGLE(fmd_matt_makeCuboidFCC, LATTICE_FCC)
CUBOID_MIX(fmd_matt_makeCuboidSC_mix, LATTICE_SC)
CUBOID_MIX(fmd_matt_makeCuboidBCC_mix, LATTICE_BCC)
CUBOID_MIX(fmd_matt_makeCuboidFCC_mix, LATTICE_FCC)

/* ------------------------------------------------- edits on resident atoms */

/* download, edit and re-upload; the calls below are rare set-up operations */
typedef void (*atom_edit_fn)(fmd_t *md, unsigned n, double *x, double *v, int *grp, unsigned *kind, void *arg);

static void edit_device_atoms(fmd_t *md, atom_edit_fn fn, void *arg)
{
    unsigned n, *aid, *kind, *cnt;
    double *x, *v;
    int *grp;
    if (!fmdh_download_all(md, &n, &x, &v, &grp, &aid, &kind, &cnt)) return;
    fn(md, n, x, v, grp, kind, arg);
    if (fmdx_upload(md->dev, n, cnt, x, v, grp, aid, kind) != FMDX_OK)
        fmdh_err_function_failed(md, true, HERE, "fmdx_upload");
    free(x); free(v); free(grp); free(aid); free(kind); free(cnt);
}

void fmd_matt_addVelocity(fmd_t *md, int GroupID, fmd_real_t vx, fmd_real_t vy, fmd_real_t vz)
{
    if (md->ParticlesDistributed)
    {
        if (fmdx_add_velocity(md->dev, GroupID, vx, vy, vz) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_add_velocity");
    }
    else if (md->gg)
    {
        for (unsigned ic = 0; ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++)
                if (GroupID == FMD_GROUP_ALL || GroupID == c->grp[i])
                {
                    c->v[3 * i] += vx;
                    c->v[3 * i + 1] += vy;
                    c->v[3 * i + 2] += vz;
                }
        }
    }
    if (md->ActiveGroup == FMD_GROUP_ALL || GroupID == FMD_GROUP_ALL || md->ActiveGroup == GroupID)
    {
        md->KineticEnergyUpdated = false;
        md->thermo_host_valid = false;
    }
}

void fmd_matt_findLimits(fmd_t *md, int GroupID, fmd_rtriple_t LowerLimit, fmd_rtriple_t UpperLimit)
{
    for (int d = 0; d < 3; d++)
    {
        LowerLimit[d] = DBL_MAX;
        UpperLimit[d] = DBL_MIN;
    }
    if (md->ParticlesDistributed)
    {
        if (fmdx_find_limits(md->dev, GroupID, LowerLimit, UpperLimit) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_find_limits");
        return;
    }
    if (!md->gg) return;
    for (unsigned ic = 0; ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        for (unsigned i = 0; i < c->n; i++)
            if (GroupID == FMD_GROUP_ALL || GroupID == c->grp[i])
                for (int d = 0; d < 3; d++)
                {
                    if (c->x[3 * i + d] < LowerLimit[d]) LowerLimit[d] = c->x[3 * i + d];
                    if (c->x[3 * i + d] > UpperLimit[d]) UpperLimit[d] = c->x[3 * i + d];
                }
    }
}

static void edit_group(fmd_t *md, unsigned n, double *x, double *v, int *grp, unsigned *kind, void *arg)
{
    (void)md; (void)x; (void)v; (void)kind;
    int *oldnew = arg;
    for (unsigned i = 0; i < n; i++)
        if (oldnew[0] == FMD_GROUP_ALL || grp[i] == oldnew[0]) grp[i] = oldnew[1];
}

void fmd_matt_changeGroupID(fmd_t *md, int old, int new)
{
    if (new < 0)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "GroupID", new);
        return;
    }
    if (md->ParticlesDistributed)
    {
        int on[2] = {old, new};
        edit_device_atoms(md, edit_group, on);
    }
    else if (md->gg)
    {
        for (unsigned ic = 0; ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++)
                if (old == FMD_GROUP_ALL || c->grp[i] == old) c->grp[i] = new;
        }
    }
    if (md->ActiveGroup != FMD_GROUP_ALL && (md->ActiveGroup == old || md->ActiveGroup == new))
    {
        md->KineticEnergyUpdated = false;
        md->thermo_host_valid = false;
    }
}

struct maxwell_arg
{
    int group;
    double temp;
};

static void edit_maxwell(fmd_t *md, unsigned n, double *x, double *v, int *grp, unsigned *kind, void *arg)
{
    (void)x;
    struct maxwell_arg *a = arg;
    for (unsigned i = 0; i < n; i++)
        if (a->group == FMD_GROUP_ALL || a->group == grp[i])
        {
            double sd = sqrt(K_BOLTZMANN * a->temp / md->kinds[kind[i]].mass);
            for (int d = 0; d < 3; d++) v[3 * i + d] = fmdh_rng_gaussian(md, sd);
        }
}

void fmd_matt_giveMaxwellDistribution(fmd_t *md, int GroupID, fmd_real_t temp)
{
    struct maxwell_arg a = {GroupID, temp};
    if (md->ParticlesDistributed)
        edit_device_atoms(md, edit_maxwell, &a);
    else if (md->gg)
    {
        for (unsigned ic = 0; ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++)
                if (GroupID == FMD_GROUP_ALL || GroupID == c->grp[i])
                {
                    double sd = sqrt(K_BOLTZMANN * temp / md->kinds[c->kind[i]].mass);
                    for (int d = 0; d < 3; d++) c->v[3 * i + d] = fmdh_rng_gaussian(md, sd);
                }
        }
    }
    if (md->ActiveGroup == FMD_GROUP_ALL || GroupID == FMD_GROUP_ALL || md->ActiveGroup == GroupID)
    {
        md->KineticEnergyUpdated = false;
        md->thermo_host_valid = false;
    }
}

/* fmd_matt_translate, matter.c:496-552 (only before distribution, as in the reference) */
void fmd_matt_translate(fmd_t *md, int GroupID, fmd_real_t dx, fmd_real_t dy, fmd_real_t dz)
{
    if (md->ParticlesDistributed)
    {
        fmdh_err_not_supported(md, false, HERE,
                               "The current version cannot translate matter after particles are distributed among subdomains.");
        return;
    }
    if (md->ActiveGroup == FMD_GROUP_ALL || GroupID == FMD_GROUP_ALL || md->ActiveGroup == GroupID)
        if ((!md->PBC[0] && dx != 0.) || (!md->PBC[1] && dy != 0.) || (!md->PBC[2] && dz != 0.))
            md->KineticEnergyUpdated = false;
    if (!md->gg) return;
    double dr[3] = {dx, dy, dz};
    for (unsigned ic = 0; ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        for (int i = 0; i < (int)c->n; i++)
            if (GroupID == FMD_GROUP_ALL || GroupID == c->grp[i])
                for (int d = 0; d < 3; d++)
                {
                    double *p = &c->x[3 * i + d];
                    *p += dr[d];
                    if (*p < 0.0 || *p >= md->l[d])
                    {
                        if (!md->PBC[d])
                        {
                            fmdh_cell_remove(c, (unsigned)i);
                            md->TotalNoOfParticles--;
                            i--;
                            break;
                        }
                        else
                            *p += (*p < 0.0 ? md->l[d] : -md->l[d]);
                    }
                }
    }
    for (unsigned ic = 0; ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        for (int i = 0; i < (int)c->n; i++)
            if (GroupID == FMD_GROUP_ALL || GroupID == c->grp[i])
            {
                int j[3];
                for (int d = 0; d < 3; d++)
                {
                    j[d] = (int)floor(c->x[3 * i + d] / md->cellh[d]);
                    if (j[d] < 0 || j[d] >= md->nc[d])
                    {
                        fmdh_err_unexpected_position(md, false, HERE);
                        return;
                    }
                }
                hcell_t *c2 = &md->gg[flat3(md->nc, j)];
                if (c != c2)
                {
                    unsigned k = fmdh_cell_new(md, c2);
                    cell_copy(c, (unsigned)i, c2, k);
                    fmdh_cell_remove(c, (unsigned)i);
                    i--;
                }
            }
    }
}

/* ------------------------------------------------------ group quantities */

/* compute_GroupTemperature_etc_globalgrid, matter.c:70-115 */
static void thermo_globalgrid(fmd_t *md)
{
    md->GroupParticlesNum = 0;
    md->GroupMomentum[0] = md->GroupMomentum[1] = md->GroupMomentum[2] = 0.;
    double mv2 = 0.;
    if (md->gg)
        for (unsigned ic = 0; ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++)
            {
                if (md->ActiveGroup != FMD_GROUP_ALL && c->grp[i] != md->ActiveGroup) continue;
                md->GroupParticlesNum++;
                double m = md->kinds[c->kind[i]].mass;
                for (int d = 0; d < 3; d++)
                {
                    mv2 += m * (c->v[3 * i + d] * c->v[3 * i + d]);
                    md->GroupMomentum[d] += m * c->v[3 * i + d];
                }
            }
        }
    md->GroupTemperature = mv2 / (3.0 * md->GroupParticlesNum * K_BOLTZMANN);
    md->GroupKineticEnergy = 3.0 / 2.0 * md->GroupParticlesNum * K_BOLTZMANN * md->GroupTemperature;
    md->KineticEnergyUpdated = true;
    md->thermo_host_valid = true;
}

static void ensure_thermo(fmd_t *md)
{
    if (!md->ParticlesDistributed)
    {
        if (!md->KineticEnergyUpdated) thermo_globalgrid(md);
        return;
    }
    if (!md->KineticEnergyUpdated)
    {
        if (fmdx_thermo(md->dev, md->ActiveGroup) != FMDX_OK)
            fmdh_err_function_failed(md, true, HERE, "fmdx_thermo");
        md->KineticEnergyUpdated = true;
        md->thermo_host_valid = false;
    }
    fmdh_sync_thermo(md);
}

fmd_real_t fmd_matt_getKineticEnergy(fmd_t *md)
{
    ensure_thermo(md);
    return md->GroupKineticEnergy;
}

fmd_real_t fmd_matt_getTotalEnergy(fmd_t *md)
{
    ensure_thermo(md);
    return md->GroupKineticEnergy + md->GroupPotentialEnergy;
}

fmd_real_t fmd_matt_getTemperature(fmd_t *md)
{
    ensure_thermo(md);
    return md->GroupTemperature;
}

void fmd_matt_getMomentum(fmd_t *md, fmd_rtriple_t out)
{
    ensure_thermo(md);
    for (int d = 0; d < 3; d++) out[d] = md->GroupMomentum[d] * MD_MASS_UNIT;
}

/* ------------------------------------------------------- configurations */

typedef struct
{
    float x[3];
    float var;
    unsigned atomkind;
} config_atom_t;

static void save_xyz(fmd_t *md, const config_atom_t *g, unsigned n)
{
    fprintf(md->ConfigFilep, "%u\n\n", n);
    for (unsigned i = 0; i < md->nkinds; i++)
        for (unsigned k = 0; k < n; k++)
            if (g[k].atomkind == i)
                fprintf(md->ConfigFilep, "%s\t%.2f\t%.2f\t%.2f\n", md->kinds[i].name, g[k].x[0], g[k].x[1], g[k].x[2]);
}

static void save_vtf(fmd_t *md, const config_atom_t *g, unsigned n)
{
    char path[MAX_PATH_LENGTH + 64];
    snprintf(path, sizeof path, "%s%05d.vtf", md->SaveDirectory, md->_FileIndex++);
    FILE *fp = fopen(path, "w");
    if (!fp)
    {
        fmdh_err_unable_open_file(md, false, HERE, path);
        return;
    }
    unsigned id = 0;
    for (unsigned i = 0; i < md->nkinds; i++)
    {
        unsigned cnt = 0;
        for (unsigned k = 0; k < n; k++)
            if (g[k].atomkind == i) cnt++;
        fprintf(fp, "atom %u:%d\tname %s\n", id, (int)(id + cnt) - 1, md->kinds[i].name);
        id += cnt;
    }
    id = 0;
    for (unsigned i = 0; i < md->nkinds; i++)
        for (unsigned k = 0; k < n; k++)
            if (g[k].atomkind == i) fprintf(fp, "%u\tbeta %.4f\n", id++, g[k].var);
    fprintf(fp, "timestep\n");
    for (unsigned i = 0; i < md->nkinds; i++)
        for (unsigned k = 0; k < n; k++)
            if (g[k].atomkind == i) fprintf(fp, "%.2f\t%.2f\t%.2f\n", g[k].x[0], g[k].x[1], g[k].x[2]);
    fclose(fp);
}

void fmd_matt_saveConfiguration(fmd_t *md)
{
    unsigned n = 0;
    config_atom_t *g = NULL;
    if (md->ParticlesDistributed)
    {
        unsigned *aid, *kind, *cnt;
        double *x, *v;
        int *grp;
        if (!fmdh_download_all(md, &n, &x, &v, &grp, &aid, &kind, &cnt)) return;
        g = malloc((n ? n : 1) * sizeof *g);
        for (unsigned i = 0; i < n; i++)
        {
            for (int d = 0; d < 3; d++) g[i].x[d] = (float)x[3 * i + d];
            g[i].var = (float)grp[i];
            g[i].atomkind = kind[i];
        }
        free(x); free(v); free(grp); free(aid); free(kind); free(cnt);
    }
    else
    {
        for (unsigned ic = 0; md->gg && ic < md->ngg; ic++) n += md->gg[ic].n;
        g = malloc((n ? n : 1) * sizeof *g);
        unsigned k = 0;
        for (unsigned ic = 0; md->gg && ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++, k++)
            {
                for (int d = 0; d < 3; d++) g[k].x[d] = (float)c->x[3 * i + d];
                g[k].var = (float)c->grp[i];
                g[k].atomkind = c->kind[i];
            }
        }
    }
    md->TotalNoOfParticles = n;
    char path[MAX_PATH_LENGTH + 64];
    switch (md->SaveConfigMode)
    {
        case FMD_SCM_XYZ_ATOMSNUM:
            if ((int)n != md->_OldNumberOfParticles)
            {
                if (md->_OldNumberOfParticles != -1 && md->ConfigFilep) fclose(md->ConfigFilep);
                snprintf(path, sizeof path, "%s%u.xyz", md->SaveDirectory, n);
                md->ConfigFilep = fopen(path, "w");
                if (!md->ConfigFilep)
                {
                    fmdh_err_unable_open_file(md, false, HERE, path);
                    md->_OldNumberOfParticles = -1;
                    break;
                }
                md->_OldNumberOfParticles = (int)n;
            }
            save_xyz(md, g, n);
            fflush(md->ConfigFilep);
            break;
        case FMD_SCM_XYZ_SEPARATE:
            snprintf(path, sizeof path, "%s%05d.xyz", md->SaveDirectory, md->_FileIndex++);
            {
                FILE *old = md->ConfigFilep;
                md->ConfigFilep = fopen(path, "w");
                if (md->ConfigFilep)
                {
                    save_xyz(md, g, n);
                    fclose(md->ConfigFilep);
                }
                else
                    fmdh_err_unable_open_file(md, false, HERE, path);
                md->ConfigFilep = old;
            }
            break;
        case FMD_SCM_CSV:
            snprintf(path, sizeof path, "%s%05d.csv", md->SaveDirectory, md->_FileIndex++);
            {
                FILE *fp = fopen(path, "w");
                if (!fp)
                {
                    fmdh_err_unable_open_file(md, false, HERE, path);
                    break;
                }
                for (unsigned i = 0; i < md->nkinds; i++)
                    for (unsigned k = 0; k < n; k++)
                        if (g[k].atomkind == i)
                            fprintf(fp, "%.2f, %.2f, %.2f, %u, %.4f\n", g[k].x[0], g[k].x[1], g[k].x[2], i, g[k].var);
                fclose(fp);
            }
            break;
        case FMD_SCM_VTF:
            save_vtf(md, g, n);
            break;
    }
    free(g);
}
