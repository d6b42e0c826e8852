/*
 * fmdx_noalloc.hip -- fmdx_update_forces_noalloc: the force evaluation of
 * fmdx_update_forces with the call's scratch in one fixed device array instead
 * of the system's buffers, for callers that cannot allocate device memory.
 * No counterpart in the reference.
 *
 * Scratch of one call (all written before they are read within the call, so
 * nothing needs clearing between calls):
 *   epart     energy partial sums           2 * NPART_MAX doubles
 *   cell_pe   pair energy per home cell     ncell doubles
 *   cell_emb  embedding energy per cell     ncell doubles
 *   htile     tile sizes per density round  ncell u64
 *   htrips    hit words recorded per cell   ncell ints (always -1 here)
 * The density pass keeps no hit words (hbits is null), so the force pass
 * rescans every cell; the words it finds are the ones the density pass would
 * have stored, and the forces are the same bits as fmdx_update_forces gives.
 * Every call on one GPU uses that GPU's copy of the array: calls that overlap
 * on one GPU race.
 */
#include <mutex>
#include "fmdx_sys.h"

#define NOALLOC_BYTES (16u << 20)
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

#define MAX_DEVICES 64
static void *scratch_dev[MAX_DEVICES];   /* device address of `scratch` on each GPU */
static std::mutex scratch_mu;

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

/* bytes of scratch a system of ncell cells takes */
static size_t scratch_bytes(long ncell)
{
    return up256(sizeof(double) * 2 * NPART_MAX) + 2 * up256(sizeof(double) * ncell) +
           up256(sizeof(unsigned long long) * ncell) + up256(sizeof(int) * ncell);
}

/* the address of `scratch` on the current device, looked up once per device; the array's name
   in host code is only a stand-in for the symbol and is never passed to a kernel */
static int scratch_address(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= MAX_DEVICES) return FMDX_EARG;
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_dev[dev])
    {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FMDX_EHIP;
        scratch_dev[dev] = p;
    }
    *out = (unsigned char *)scratch_dev[dev];
    return FMDX_OK;
}

extern "C" int fmdx_update_forces_noalloc(fmdx_system_t *s, int active_group, int ttm_friction)
{
    static_assert(FMDX_NOALLOC_MAX_CELLS <= 600000, "scratch layout");
    if (!s) return FMDX_EARG;
    if (s->D.ncell > FMDX_NOALLOC_MAX_CELLS || scratch_bytes(s->D.ncell) > NOALLOC_BYTES) return FMDX_ESIZE;
    if (!s->uploaded) return FMDX_ESTATE;
    if (s->D.has_eam && !s->rhoc) return FMDX_ESTATE;
    int cur = -1;
    CK(hipGetDevice(&cur));
    if (cur != s->dev) CK(hipSetDevice(s->dev));
    unsigned char *base = nullptr;
    int r = scratch_address(s->dev, &base);
    if (r == FMDX_OK)
    {
        const long n = s->D.ncell;
        KDesc D = s->D;
        D.epart = (double *)base;
        base += up256(sizeof(double) * 2 * NPART_MAX);
        D.cell_pe = (double *)base;
        base += up256(sizeof(double) * n);
        D.cell_emb = (double *)base;
        base += up256(sizeof(double) * n);
        D.htile = (unsigned long long *)base;
        base += up256(sizeof(unsigned long long) * n);
        D.htrips = (int *)base;
        D.hbits = nullptr;
        D.ag = active_group;
        D.energy = 1;
        D.ttm = ttm_dev(s, ttm_friction);
        s->D.ag = active_group;
        if (launch_forces(D, s->st, s->precision) != hipSuccess) r = FMDX_EHIP;
    }
    if (cur != s->dev) hipSetDevice(cur);
    return r;
}
