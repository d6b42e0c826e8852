/*
 * fmdx_internal.h -- device-side layout shared by the gfx950 kernels.
 *
 * HBM layout of one system (all arrays SoA):
 *   cells   : interior cells only, flat index c = (ix*ny + iy)*nz + iz, the
 *             reference's interior storage order (subdomain.c:201-222, ns=1).
 *             There is no ghost layer: periodic images are formed inside the
 *             kernels by shifting positions by the box length.
 *   slots   : cell c owns slots [c*cap, c*cap + cnt[c]); the within-cell order
 *             of those slots is the reference's within-cell atom order.
 *   per slot: x y z vx vy vz fx fy fz aux (f64), xq (f32 x y z cell-relative
 *             and F', mixed build), grp (i32), aid (u32), kind (u8), mv (i32 binning
 *             destination during a re-binning, the atom's turi index during
 *             the TTM field sums), code (u8 migration code)
 */
#ifndef FMDX_INTERNAL_H
#define FMDX_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fmdx.h"

#define KB_EV 8.6173303e-5     /* general.h:46 */
#define MAXK FMDX_MAX_KINDS
#define MAXE FMDX_MAX_ELEMENTS
#define NPART_MAX 4096         /* blocks reduced by the single-block reductions */
#define HBITS_WORDS 4          /* 64-bit hit words per lane kept between the EAM passes */
#define NCOUNTERS 16           /* status words per system */

/* mover destination codes */
#define MV_STAY (-1)
#define MV_DELETE (-2)

struct PairParam
{
    int cat;
    int iloc, jloc;
    int pad;
    double rc2;
    double p0, p1, p2;
};

/* per-interval cubic coefficients of a reference natural spline:
   value = c0 + b*(c1 + b*(c2 + b*c3)),  b = x/h - k  (cspline.h:26-27 rewritten in b) */
struct EamDev
{
    int ne, nrho, nr2;
    double drho, inv_drho, dr2, inv_dr2, rc2;
    double rx2;            /* FP32 kernels: pairs from this r^2 up are evaluated in FP64 (k_force.hip) */
    const double4 *Fc;     /* [ne][nrho-1] */
    const double4 *rhoc;   /* [ne][nr2-1]  */
    const double4 *phic;   /* [ne*ne][nr2-1] */
    const float4 *Fc32, *rhoc32, *phic32;
    /* element 0's force record per interval (k_force.hip, FRec): FP64 2 x double4,
       (rho1 rho2 rho3 phi1)(phi2 phi3 rho0 phi0); FP32 2 x float4, same order */
    const double4 *rf;
    const float4 *rf32;
    /* element 0's derivative knots, FP32, k < nr2: (rho D_k, phi D_k, rho Y_k, phi Y_k) with
       D_k = h y'(x_k) and Y_k = y''_k h^2/6: the table the LDS-resident force pass holds
       (k_force.hip, FMDX_EAM_LDS) */
    const float4 *kn32;
};

struct TtmDev
{
    int on;
    int ncz;
    double tcellh_z;
    const double *xi;      /* [ncx*ncy*(ncz+2)] */
    const double *vcm;     /* [3*ncx*ncy*(ncz+2)] */
    int ncx, ncy;          /* a turi that is not 1 x 1 x n: friction on every component */
    double tcellh_x, tcellh_y;
};

struct KDesc
{
    /* geometry */
    int nc0, nc1, nc2, ncell;
    int cap;
    int pbc0, pbc1, pbc2;
    double L0, L1, L2;
    double h0, h1, h2;
    /* slots */
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *aux;
    float4 *xq;   /* mixed build: x y z relative to the home cell origin (FP32), and F'(rho) in w */
    float *np;    /* mixed build, single-kind LJ or Morse, each pair once: partner sums
                     [3][13][ncell*cap] (x, y, z; forward direction; slot), or null */
    int xq_keep;  /* mixed build: the integrators write xq and the re-binning carries it along */
    char *slot_mem;   /* the one allocation the per-slot arrays, mover staging and part are carved from */
    int *grp;
    unsigned *aid;
    uint8_t *kind;
    int *cnt;
    /* binning scratch */
    int *mv;
    uint8_t *code;
    int *nleave, *narr, *nearly, *fill;
    int *btouch;          /* [ncell/TPB + 1] a re-binning moves atoms out of or into block b of
                             TPB cells (set by the start step, cleared by the batch's kb_touch) */
    unsigned long long *bkey;   /* [ncell*cap] arrival buckets */
    int *bid;
    double *mvx, *mvy, *mvz, *mvvx, *mvvy, *mvvz;   /* mover staging */
    int *mvgrp;
    unsigned *mvaid;
    uint8_t *mvkind;
    unsigned *counters;   /* [0] movers, [1] deleted, [2] max cell count, [3] error word, [4] cell + 1 of
                             the first non-finite force, [5] cell + 1 of the first unbinnable position,
                             [6] force-pass rescans after a tile mismatch, [7] exact-shell pairs not
                             found (evaluated in FP32), [8..10] audit: atoms, fullest
                             cell, atoms with a non-finite x, v or F (NCOUNTERS entries) */
    /* potentials */
    int nkinds;
    int single_cat;        /* category when one category is used everywhere, else -1 */
    int has_eam;
    double rcmax2, rceam2;
    double mass[MAXK];
    int elem[MAXK];
    const PairParam *pairs;   /* [MAXK*MAXK] */
    EamDev eam;
    /* pair-candidate hits found by the density pass, replayed by the force pass */
    unsigned long long *hbits;   /* [ncell][HBITS_WORDS][threads per cell] one bit per (thread, scan trip) */
    int *htrips;                 /* [ncell] hit words recorded, -1 = none (the force pass rescans) */
    unsigned long long *htile;   /* [ncell] candidates staged in each of the first 4 tile rounds, 16 bits each */
    /* reductions */
    double *cell_pe;       /* [ncell] pair energy per home cell */
    double *cell_emb;      /* [ncell] embedding energy per home cell */
    double *part;          /* [blocks*8] per-block thermo partials */
    double *epart;         /* [NPART_MAX*2] per-block energy partials */
    double *thermo;        /* fmdx_thermo_t mirror, doubles first */
    /* step parameters, set by the host before a launch */
    int ag;                /* active group, -1 = every atom */
    int thermostat;        /* Berendsen rescale in the start step */
    int energy;            /* the force call also sums the potential energy */
    double dt, tau, T0;
    double hdtm[MAXK];     /* dt * 0.5 / mass per kind (integrators.c:67) */
    TtmDev ttm;            /* TTM friction source (forces.c:292-330) */
};

/* FP32 position relative to the origin of cell (cx, cy, cz), formed in FP64 without FMA
   contraction and rounded once: the mixed build's force kernels read it (xq x y z; w is
   F'(rho), written by the density pass).  The integrators, the re-binning and k_relpos all
   form it here, so that each gives the same bits. */
static __device__ __forceinline__ float4 rel_pos(const KDesc &D, int cx, int cy, int cz, double x, double y,
                                                 double z)
{
#pragma clang fp contract(off)
    const double ox = cx * D.h0, oy = cy * D.h1, oz = cz * D.h2;
    return make_float4((float)(x - ox), (float)(y - oy), (float)(z - oz), 0.0f);
}

/* many systems in one launch: block b works on block items[b].y of system items[b].x.
   With `count` set, the list was compacted on the device and holds *count items (see
   each_item); `skip` items at its head are another launch's. */
struct BatchMap
{
    const KDesc *descs;
    const int2 *items;
    const int *count;
    int skip;
};

/* f(descriptor, block) for the items of the workgroup.  A full list: the workgroup's own
   item.  A compacted list: the items from `skip` on are split into one contiguous run per
   XCD (workgroup b runs on XCD b % 8, so that each XCD's L2 holds the neighbours of
   nearby cells); LOOP: the workgroup takes every (gridDim/8)-th item of its XCD's run;
   otherwise it takes at most one, and the items beyond gridDim are left to a LOOP launch
   with skip = gridDim. */
template <bool LOOP, class F>
static __device__ __forceinline__ void each_item_xy(const BatchMap &M, F f)
{
    int base = (int)blockIdx.x, first = 0, len = 1, step = 1;
    if (M.count)
    {
        int n = *M.count - M.skip;
        if (!LOOP) n = n < (int)gridDim.x ? n : (int)gridDim.x;
        const int x = (int)(blockIdx.x & 7), q = n >> 3, r = n & 7;
        base = M.skip + x * q + (x < r ? x : r);
        len = q + (x < r ? 1 : 0);
        first = (int)(blockIdx.x >> 3);
        step = (int)(gridDim.x >> 3);
    }
    if (!LOOP)
    {
        if (first >= len) return;
        f(M.items[base + first]);
        return;
    }
    for (int i = first; i < len; i += step)
    {
        f(M.items[base + i]);
        __syncthreads();   /* the next item reuses the workgroup's LDS */
    }
}

template <bool LOOP, class F>
static __device__ __forceinline__ void each_item(const BatchMap &M, F f)
{
    each_item_xy<LOOP>(M, [&](int2 it) { f(M.descs[it.x], it.y); });
}

/* TTM state resident on the device: n = ncx*ncy*(ncz+2) turi cells, index
   (ix*ncy + iy)*(ncz+2) + iz, ghosts in z only; dim3 selects the grid solver */
struct TtmState
{
    int type, ncz, n, ratio;
    int ncx, ncy, nzg, dim3, kmode;
    double tcellh_x, tcellh_y;
    double tcellh_z, tvol;
    double C_gamma, K, G, Kzh_A, Kzh_B, Kzh_v2;
    int nC, nG;
    const double *T_C, *Ct, *C_DD, *T_G, *Gt, *G_DD;
    double laser_factor, laser_absdepth_inv, laser_2sig2_inv, laser_t0, laser_tdiff;
    double frontsurf;
    unsigned min_atoms;
    double *mass, *vcm, *Ti, *Te, *Te2, *xi, *Ke, *Ce, *Gc;
    unsigned *num;
};

/* thermo record layout in device memory (doubles) */
enum
{
    TH_PE = 0, TH_EMB, TH_MV2, TH_KE, TH_T, TH_PX, TH_PY, TH_PZ, TH_N, TH_NDOUBLES
};

static __device__ __forceinline__ void set_err(const KDesc &D, unsigned bits)
{
    atomicOr(&D.counters[3], bits);
}

/* keep the first value written into a diagnostic word that starts at 0 */
static __device__ __forceinline__ void note_first(unsigned *w, unsigned v)
{
    atomicCAS(w, 0u, v);
}

/* The O(N) kernels map a block of TPB = 256 threads onto CPB = 16 consecutive
   cells, 16 lanes per cell; lane l visits the occupied slots l, l+16, ... of
   its cell, so only occupied slots are touched and each cell's slots are read
   as one contiguous segment. */
#define TPB 256
#define CPB 16
#define LPC (TPB / CPB)

template <class F>
static __device__ __forceinline__ void each_slot(const KDesc &D, int blk, F f)
{
    const int c = blk * CPB + (int)threadIdx.x / LPC;
    if (c >= D.ncell) return;
    const int n = D.cnt[c];
    for (int k = (int)threadIdx.x % LPC; k < n; k += LPC) f((long)c * D.cap + k, c);
}

/* blocks of the first level of the fixed-order energy sum of a system */
static __host__ __device__ inline int pe_blocks(int ncell)
{
    const int nb = (ncell + 1023) / 1024;
    return nb < 1 ? 1 : (nb > NPART_MAX ? NPART_MAX : nb);
}

/* work lists of a batch of systems (device arrays of (system, logical block)) */
struct BatchPlan
{
    const KDesc *descs;          /* [nsys] */
    int nsys;
    const int2 *slot_items;      /* (system, group of CPB cells): O(N) kernels */
    int n_slot;
    const int2 *cellblk_items;   /* (system, block of TPB cells): binning */
    int n_cellblk;
    const int2 *pe_items;        /* (system, energy partial block) */
    int n_pe;
    const int2 *cell0[4];        /* (system, cell) for the force pass, per pair category */
    int n_cell0[4];
    const int2 *cell1[4];        /* (system, cell) for the EAM density pass, per pair category */
    int n_cell1[4];
    int max_nkinds;
    int n3;                      /* the LJ and Morse members evaluate each pair once (all have partner sums) */
    /* the step over occupied cells only: after each re-binning kb_occ writes the lists below
       and their lengths occ_n[0] (slot groups), [1 + cat] (force pass), [5 + cat] (density pass);
       before each re-binning kb_touch lists the blocks of TPB cells it changes, occ_n[9];
       occ_n[10]: every occupied cell, for the start step */
    int compact;
    /* grids that walk the compacted lists (multiples of 8), sized by the host from the list
       lengths it last read, with room to grow; a list that outgrows its grid is still walked
       whole, some workgroups taking more than one item */
    int g_slot, g_cell0[4], g_cell1[4], g_bin, g_all;
    const int *sys_cat;          /* [nsys] pair category of each member's force pass */
    int *occ_n;                  /* [16] */
    int2 *occ_slot;              /* capacity n_slot */
    int2 *occ_cell0[4];          /* capacity n_cell0[cat] */
    int2 *occ_cell1[4];          /* capacity n_cell1[cat] */
    int2 *occ_bin;               /* capacity n_cellblk */
    int2 *occ_all;               /* capacity the sum of n_cell0 */
};

/* kernel launchers (defined in the kernel translation units); the step
   parameters (active group, dt, thermostat, TTM friction) travel in KDesc */
hipError_t launch_vv_start(const KDesc &D, hipStream_t st);
hipError_t launch_vv_finish(const KDesc &D, hipStream_t st);
hipError_t launch_se_step(const KDesc &D, hipStream_t st);
hipError_t launch_thermo(const KDesc &D, hipStream_t st);
hipError_t launch_flag_only(const KDesc &D, hipStream_t st);
hipError_t launch_bin(const KDesc &D, hipStream_t st);
hipError_t launch_reduce_pe(const KDesc &D, hipStream_t st);
/* relpos: form the FP32 cell-relative positions first (mixed build), unless the
   last integrator step and re-binning left them in place */
hipError_t launch_forces(const KDesc &D, hipStream_t st, int precision, bool relpos = true);
hipError_t launch_pass1(const KDesc &D, hipStream_t st, int precision, bool relpos = true);
hipError_t launch_pass0(const KDesc &D, hipStream_t st, int precision);
hipError_t launchb_vv_start(const BatchPlan &P, hipStream_t st);
hipError_t launchb_vv_finish(const BatchPlan &P, hipStream_t st);
hipError_t launchb_thermo(const BatchPlan &P, hipStream_t st);
hipError_t launchb_reduce_pe(const BatchPlan &P, hipStream_t st);
/* energy: also the pair energies and each member's PE sum (steps whose status is read) */
hipError_t launchb_forces(const BatchPlan &P, hipStream_t st, int precision, int energy, bool relpos = true);
/* each pair evaluated once (partner sums, then a fold) for this system in this precision */
bool n3_wanted(const KDesc &D, int precision);
hipError_t launchb_collect(const BatchPlan &P, hipStream_t st, unsigned *status, double *thermo);
/* compact: the occupied-cell lists from the members' current cell counts */
hipError_t launchb_occ(const BatchPlan &P, hipStream_t st);
hipError_t launch_add_velocity(const KDesc &D, hipStream_t st, int group, double vx, double vy, double vz);
hipError_t launch_relayout(const KDesc &Dold, const KDesc &Dnew, hipStream_t st);
hipError_t launch_scatter(const KDesc &D, hipStream_t st, const int *cell_start, const double *x, const double *v,
                          const int *grp, const unsigned *aid, const uint8_t *kind);
hipError_t launch_gather(const KDesc &D, hipStream_t st, const int *cell_start, double *x, double *v, double *f,
                         double *aux, int *grp, unsigned *aid, uint8_t *kind);
hipError_t launch_ttm_fields(const KDesc &D, const TtmState &S, hipStream_t st);
hipError_t launch_ttm_presolve(const TtmState &S, unsigned *err, hipStream_t st);
hipError_t launch_ttm_solve(const TtmState &S, double time, double dt, unsigned *err, hipStream_t st);
hipError_t launch_fill(double *p, long n, double v, hipStream_t st);
hipError_t launch_audit(const KDesc &D, hipStream_t st);
/* threads per home cell of the pair kernels (the hit words are kept per thread); k_force.hip */
int force_threads_per_cell();

#endif /* FMDX_INTERNAL_H */
