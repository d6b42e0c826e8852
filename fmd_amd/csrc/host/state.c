/*
 * state.c -- text state files (checkpoint / resume and the injection format
 * for identical inputs).
 *
 * Format and order follow fmd_io_saveState / fmd_io_loadState
 * (misc.c:121-266, 310-422): time, atom count, box, PBC flags, then for each
 * atom in cell order "name group", "x y z", "vx vy vz" at %.16e, which
 * round-trips doubles exactly.  Loading appends atoms to the global grid in
 * file order, binned by floor(x / cellh), so saved and reloaded states have
 * the same within-cell order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_host.h"

static const char fmt3[] = "%.16e\t%.16e\t%.16e\n";

void fmd_io_loadState(fmd_t *md, fmd_string_t path, bool UseTime)
{
    if (md->ParticlesDistributed)
    {
        fmdh_err_not_supported(md, false, HERE, "A state cannot be loaded after the particles are distributed.");
        return;
    }
    FILE *fp = fopen(path, "r");
    if (!fp)
    {
        fmdh_err_unable_open_file(md, false, HERE, path);
        return;
    }
    double t, l0, l1, l2;
    unsigned n;
    int p0, p1, p2;
    if (fscanf(fp, "%lf", &t) != 1 || fscanf(fp, "%u\n", &n) != 1 || fscanf(fp, "%lf%lf%lf", &l0, &l1, &l2) != 3 ||
        fscanf(fp, "%d%d%d", &p0, &p1, &p2) != 3)
    {
        fclose(fp);
        fmdh_err_file_corrupted(md, false, HERE, "state", path);
        return;
    }
    if (UseTime) md->time = t;
    if (!md->BoxSizeDetermined)
    {
        md->l[0] = l0; md->l[1] = l1; md->l[2] = l2;
        md->BoxSizeDetermined = true;
    }
    if (!md->PBCdetermined)
    {
        md->PBC[0] = p0; md->PBC[1] = p1; md->PBC[2] = p2;
        md->PBCdetermined = true;
    }
    if (!md->grid_made && !fmdh_make_grid(md))
    {
        fclose(fp);
        return;
    }
    for (unsigned i = 0; i < n; i++)
    {
        char name[64];
        int g;
        double x[3], v[3];
        if (fscanf(fp, "%63s%d", name, &g) != 2)
        {
            fclose(fp);
            fmdh_err_file_corrupted(md, false, HERE, "state", path);
            return;
        }
        if (g < 0)
        {
            fclose(fp);
            fmdh_err_unacceptable_int(md, false, HERE, "GroupID", g);
            return;
        }
        unsigned k;
        for (k = 0; k < md->nkinds; k++)
            if (strcmp(name, md->kinds[k].name) == 0) break;
        if (k == md->nkinds)
        {
            fclose(fp);
            fmdh_err_unacceptable_int(md, false, HERE, "atomkind", (int)k);
            return;
        }
        if (fscanf(fp, "%lf%lf%lf", &x[0], &x[1], &x[2]) != 3 || fscanf(fp, "%lf%lf%lf", &v[0], &v[1], &v[2]) != 3)
        {
            fclose(fp);
            fmdh_err_file_corrupted(md, false, HERE, "state", path);
            return;
        }
        int ic[3];
        for (int d = 0; d < 3; d++)
        {
            ic[d] = (int)floor(x[d] / md->cellh[d]);
            if (ic[d] < 0 || ic[d] >= md->nc[d])
            {
                fclose(fp);
                fmdh_err_unexpected_position(md, false, HERE);
                return;
            }
        }
        hcell_t *c = &md->gg[ic[2] + md->nc[2] * (ic[1] + md->nc[1] * ic[0])];
        unsigned j = fmdh_cell_new(md, c);
        c->grp[j] = g;
        c->kind[j] = k;
        c->aid[j] = md->TotalNoOfParticles++;
        for (int d = 0; d < 3; d++)
        {
            c->x[3 * j + d] = x[d];
            c->v[3 * j + d] = v[d];
        }
    }
    fclose(fp);
    md->KineticEnergyUpdated = false;
    md->thermo_host_valid = false;
}

void fmd_io_saveState(fmd_t *md, fmd_string_t filename)
{
    char path[MAX_PATH_LENGTH + 64];
    snprintf(path, sizeof path, "%s%s", md->SaveDirectory, filename);
    FILE *fp = fopen(path, "w");
    if (!fp)
    {
        fmdh_err_unable_open_file(md, false, HERE, path);
        return;
    }
    if (!md->ParticlesDistributed)
    {
        unsigned n = 0;
        for (unsigned ic = 0; md->gg && ic < md->ngg; ic++) n += md->gg[ic].n;
        fprintf(fp, "%.16e\n", md->time);
        fprintf(fp, "%u\n", n);
        fprintf(fp, fmt3, md->l[0], md->l[1], md->l[2]);
        fprintf(fp, "%d %d %d\n", md->PBC[0], md->PBC[1], md->PBC[2]);
        for (unsigned ic = 0; md->gg && ic < md->ngg; ic++)
        {
            hcell_t *c = &md->gg[ic];
            for (unsigned i = 0; i < c->n; i++)
            {
                fprintf(fp, "%s %d\n", md->kinds[c->kind[i]].name, c->grp[i]);
                fprintf(fp, fmt3, c->x[3 * i], c->x[3 * i + 1], c->x[3 * i + 2]);
                fprintf(fp, fmt3, c->v[3 * i], c->v[3 * i + 1], c->v[3 * i + 2]);
            }
        }
        fclose(fp);
        return;
    }
    unsigned n, *aid, *kind, *cnt;
    double *x, *v;
    int *grp;
    if (!fmdh_download_all(md, &n, &x, &v, &grp, &aid, &kind, &cnt))
    {
        fclose(fp);
        return;
    }
    md->TotalNoOfParticles = n;
    fprintf(fp, "%.16e\n", md->time);
    fprintf(fp, "%u\n", n);
    fprintf(fp, fmt3, md->l[0], md->l[1], md->l[2]);
    fprintf(fp, "%d %d %d\n", md->PBC[0], md->PBC[1], md->PBC[2]);
    for (unsigned i = 0; i < n; i++)
    {
        fprintf(fp, "%s %d\n", md->kinds[kind[i]].name, grp[i]);
        fprintf(fp, fmt3, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        fprintf(fp, fmt3, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    }
    fclose(fp);
    free(x); free(v); free(grp); free(aid); free(kind); free(cnt);
}
