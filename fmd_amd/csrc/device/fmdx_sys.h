/*
 * fmdx_sys.h -- host-side state of one device system, shared by the C-ABI
 * translation units (single-system entry points and the batch engine).
 */
#ifndef FMDX_SYS_H
#define FMDX_SYS_H

#include "fmdx_internal.h"

#define NPROF 64   /* force calls whose kernels are timed in profile mode */

struct fmdx_system
{
    int dev;
    hipStream_t st;
    int precision;
    KDesc D;
    unsigned natoms_uploaded;
    int uploaded;
    int member;            /* 1 while a batch holds this system's device pointers */
    int xq_fresh;          /* mixed: the last integrator step left the FP32 cell-relative positions in place */
    double mass[MAXK];
    PairParam hp[MAXK * MAXK];
    PairParam *dpairs;
    char *cell_mem;        /* per-cell arrays, counters, reductions and dpairs: one allocation */
    /* EAM */
    double4 *Fc, *rhoc, *phic, *rf;
    float4 *Fc32, *rhoc32, *phic32, *rf32, *kn32;
    char *tab_mem;         /* the nine EAM tables: one allocation */
    /* TTM */
    int ttm_on;
    TtmState T;
    void *ttm_mem[16];
    int nttm_mem;
    /* timing */
    hipEvent_t ev[16];
    int nev;
    int profile;
    hipEvent_t pev[NPROF * 3];
    unsigned prof_n;
};

#define CK(call)                                   \
    do                                             \
    {                                              \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return FMDX_EHIP;    \
    } while (0)

template <class T>
static inline int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return FMDX_ENOMEM;
    return FMDX_OK;
}

template <class T>
static inline void dfree(T *&p)
{
    if (p) hipFree((void *)p);
    p = nullptr;
}

/* Arrays that are allocated and released together are carved out of one device
   allocation, each on a 256-byte boundary: hipFree waits for the device and costs about
   0.1 ms, so a system's ~45 arrays took ~5 ms to release one by one.  A group is laid
   out by one function called twice with the same sizes: first over no base, which only
   sums the sizes (and leaves every pointer null), then over the allocation. */
struct Carve
{
    char *base;
    size_t off;
    explicit Carve(char *b = nullptr) : base(b), off(0) {}
    template <class T>
    void operator()(T **p, size_t n)
    {
        if (n == 0) n = 1;
        *p = base ? (T *)(base + off) : nullptr;
        off += (n * sizeof(T) + 255) & ~(size_t)255;
    }
};

template <class F>
static inline int carve_alloc(char **mem, F layout)
{
    Carve sizing;
    layout(sizing);
    *mem = nullptr;
    if (hipMalloc((void **)mem, sizing.off) != hipSuccess)
    {
        *mem = nullptr;
        return FMDX_ENOMEM;
    }
    Carve c(*mem);
    layout(c);
    return FMDX_OK;
}

/* a staging group released when the call returns, error returns included */
struct StageMem
{
    char *p = nullptr;
    ~StageMem() { dfree(p); }
};

/* releases a group and nulls its arrays */
template <class F>
static inline void carve_free(char **mem, F layout)
{
    dfree(*mem);
    Carve none;
    layout(none);
}

/* step parameters of a system; md->timestep * 0.5 / mass (integrators.c:67)
   is evaluated on the host with the same operations */
static inline void set_step(KDesc &D, const double *mass, int ag, double dt, int thermostat, double tau, double T0)
{
    D.ag = ag;
    D.dt = dt;
    D.thermostat = thermostat;
    D.tau = tau;
    D.T0 = T0;
    for (int k = 0; k < MAXK; k++) D.hdtm[k] = dt * 0.5 / mass[k];
}

/* the TTM friction source of a force call (fmdx_api.hip) */
TtmDev ttm_dev(fmdx_system_t *s, int on);

/* pair-category specialisation of the force kernels for a system (k_force.hip) */
int force_category(const KDesc &D);

/* the partner-sum arrays of a system whose force passes evaluate each pair once in this
   precision (allocated on first need; fmdx_api.hip) */
int ensure_n3(fmdx_system_t *s, int precision);

#endif /* FMDX_SYS_H */
