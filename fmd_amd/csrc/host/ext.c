/*
 * ext.c -- the extension entry points declared in include/fmd_ext.h.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fmd_ext.h"
#include "fmd_host.h"

void fmd_ext_loadAtoms(fmd_t *md, unsigned n, const double *x, const double *v, const int *group,
                       const unsigned *kind)
{
    if (md->ParticlesDistributed)
    {
        fmdh_err_not_supported(md, false, HERE, "Atoms cannot be loaded after the particles are distributed.");
        return;
    }
    if (!md->grid_made && !fmdh_make_grid(md)) return;
    for (unsigned i = 0; i < n; i++)
    {
        int ic[3];
        for (int d = 0; d < 3; d++)
        {
            ic[d] = (int)floor(x[3 * i + d] / md->cellh[d]);
            if (ic[d] < 0 || ic[d] >= md->nc[d])
            {
                fmdh_err_unexpected_position(md, false, HERE);
                return;
            }
        }
        unsigned k = kind ? kind[i] : 0;
        if (k >= md->nkinds)
        {
            fmdh_err_unacceptable_int(md, false, HERE, "atomkind", (int)k);
            return;
        }
        hcell_t *c = &md->gg[ic[2] + md->nc[2] * (ic[1] + md->nc[1] * ic[0])];
        unsigned j = fmdh_cell_new(md, c);
        c->grp[j] = group ? group[i] : 0;
        c->kind[j] = k;
        c->aid[j] = md->TotalNoOfParticles++;
        for (int d = 0; d < 3; d++)
        {
            c->x[3 * j + d] = x[3 * i + d];
            c->v[3 * j + d] = v ? v[3 * i + d] : 0.0;
        }
    }
    md->KineticEnergyUpdated = false;
    md->thermo_host_valid = false;
}

bool fmd_ext_distribute(fmd_t *md)
{
    if (!fmdh_update_potcats(md)) return false;
    if (!fmdh_distribute(md)) return false;
    if (md->dev_pot_dirty && !fmdh_push_potentials(md)) return false;
    return true;
}

fmdx_system_t *fmd_ext_getDevice(fmd_t *md) { return md->dev; }

void fmd_ext_setPrecision(fmd_t *md, int precision)
{
    md->precision = precision;
    if (md->dev && fmdx_set_precision(md->dev, precision) != FMDX_OK)
        fmdh_err_unacceptable_int(md, false, HERE, "precision", precision);
}

unsigned fmd_ext_getAtoms(fmd_t *md, double *x, double *v, double *f, int *group, unsigned *atomid,
                          unsigned *kind)
{
    if (md->ParticlesDistributed)
    {
        unsigned n = 0;
        if (fmdx_download(md->dev, &n, NULL, x, v, f, NULL, group, atomid, kind) != FMDX_OK)
        {
            fmdh_err_function_failed(md, true, HERE, "fmdx_download");
            return 0;
        }
        return n;
    }
    unsigned k = 0;
    for (unsigned ic = 0; md->gg && ic < md->ngg; ic++)
    {
        hcell_t *c = &md->gg[ic];
        for (unsigned i = 0; i < c->n; i++, k++)
        {
            for (int d = 0; d < 3; d++)
            {
                if (x) x[3 * k + d] = c->x[3 * i + d];
                if (v) v[3 * k + d] = c->v[3 * i + d];
                if (f) f[3 * k + d] = 0.0;
            }
            if (group) group[k] = c->grp[i];
            if (atomid) atomid[k] = c->aid[i];
            if (kind) kind[k] = c->kind[i];
        }
    }
    return k;
}

double fmd_ext_getPotentialEnergy(fmd_t *md)
{
    fmdh_sync_thermo(md);
    return md->GroupPotentialEnergy;
}
