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
    /* EAM */
    double4 *Fc, *rhoc, *phic, *rf;
    float4 *Fc32, *rhoc32, *phic32, *rf32, *kn32;
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
