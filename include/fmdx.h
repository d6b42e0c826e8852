/*
 * fmdx.h -- C-ABI of the gfx950 step kernels (libfmdx.so).
 *
 * The reference has no plugin/operator API for its step: the seams are
 * plain C functions called from its time loop.  Each entry point below
 * replaces one of those seams for one system resident in device memory:
 *
 *   fmdx_update_forces   _fmd_dync_updateForces          src/forces.c:332-382
 *                        (LJ lj.c:28-90, Morse morse.c:30-92, EAM eam.c:29-322,
 *                         hybrid forces.c:33-290, TTM friction forces.c:292-330)
 *   fmdx_vv_start        VelocityVerlet_startStep         src/integrators.c:40-75
 *                        + _fmd_refreshGrid                src/misc.c:268-308
 *                        + _fmd_particles_migrate          src/md-ghost.c:827-869
 *   fmdx_vv_finish       VelocityVerlet_finishStep        src/integrators.c:94-126
 *   fmdx_se_step         SymplecticEuler_takeOneStep      src/integrators.c:129-179
 *   fmdx_thermo          _fmd_compute_GroupTemperature_etc_localgrid src/matter.c:31-68
 *   fmdx_ttm_update      update_field_ttm_Te_and_xi       src/turi.c:1337-1373
 *                        (field sums turi.c:1112-1261, presolve ttm.c:194-248,
 *                         solve ttm.c:250-383)
 *   fmdx_upload          transfer_from_globalgrid_to_rank0_grid src/matter.c:329-363
 *   fmdx_download        the cell walks of fmd_io_saveState src/misc.c:310-422
 *
 * All arguments are plain C types.  Every call is asynchronous on the
 * system's own stream unless it says "synchronises".  Device-side errors
 * (the reference's major errors raised inside the step) are collected in a
 * sticky error word that fmdx_read_thermo returns.
 */
#ifndef FMDX_H
#define FMDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define FMDX_MAX_KINDS 8
#define FMDX_MAX_ELEMENTS 4
#define FMDX_TTM_MAX_CELLS 1024      /* cells along z of a 1 x 1 x n turi */
#define FMDX_TTM_MAX_TCELLS 4000     /* ncx * ncy * (ncz + 2) of a turi solved on the ncx x ncy x ncz grid */

typedef struct fmdx_system fmdx_system_t;

/* return codes */
enum
{
    FMDX_OK = 0,
    FMDX_EHIP = -1,      /* a HIP runtime call failed */
    FMDX_EARG = -2,      /* invalid argument */
    FMDX_ENOMEM = -3,    /* device allocation failed */
    FMDX_ESTATE = -4,    /* call out of order (e.g. no atoms uploaded) */
    FMDX_ENODEV = -5,    /* no GPU visible */
    FMDX_ESIZE = -6      /* beyond the fixed scratch of fmdx_update_forces_noalloc */
};

/* sticky device error bits */
enum
{
    FMDX_DEV_ERR_POSITION = 1u << 0,   /* atom beyond the one-cell ghost layer, misc.c:281-291 */
    FMDX_DEV_ERR_RHO_RANGE = 1u << 1,  /* rho beyond the F(rho) table, eam.c:168-171 */
    FMDX_DEV_ERR_OVERFLOW = 1u << 2,   /* a cell outgrew its slot capacity */
    FMDX_DEV_ERR_TTM_CE = 1u << 3,     /* Te outside the Ce table, ttm.c:163-166 */
    FMDX_DEV_ERR_TTM_G = 1u << 4,      /* Te outside the G table, ttm.c:179-182 */
    FMDX_DEV_ERR_FORCE = 1u << 5       /* a non-finite force; no counterpart in the reference, where such a
                                          force shows up one step later as misc.c:281-291 */
};

/* force-kernel arithmetic */
enum
{
    FMDX_PREC_FP64 = 0,    /* FP64 everywhere */
    FMDX_PREC_MIXED = 1    /* FP32 pair arithmetic + Kahan accumulation, FP64 integration */
};

/* pair categories: same numbering as the reference's potcat_t, src/potential.h:26-32 */
enum
{
    FMDX_POT_NONE = 0,
    FMDX_POT_LJ = 1,
    FMDX_POT_EAM = 2,
    FMDX_POT_MORSE = 3
};

typedef struct
{
    double pe;               /* pair + embedding energy of the active group (eV) */
    double embed;            /* sum of F(rho) (eV) */
    double mv2;              /* sum of m v^2 over the active group (internal units) */
    double ke;               /* 0.5 * mv2 */
    double temperature;      /* mv2 / (3 N kB) */
    double momentum[3];      /* sum of m v (internal units) */
    double n_active;         /* number of atoms in the active group */
    unsigned natoms;         /* atoms resident on the device */
    unsigned max_cell_count; /* largest cell population seen by the last re-binning */
    unsigned err;            /* sticky FMDX_DEV_ERR_* bits */
    unsigned cap;            /* slots per cell */
    /* diagnostics: cell index + 1 of the first atom seen with a non-finite force or an unbinnable
       position (0: none), and the cells whose density-pass hits were rescanned by the force pass
       because its tile differed (always 0 unless the two passes disagree) */
    unsigned bad_force_cell;
    unsigned bad_position_cell;
    unsigned replay_mismatch;
    unsigned exact_unmatched;   /* mixed build: exact-shell pairs evaluated in FP32 because their FP64
                                   positions were not found (always 0 unless the search fails) */
    /* filled by fmdx_audit: atoms counted over the cells, the fullest cell, and the atoms whose
       position, velocity or force is not finite */
    unsigned audit_atoms;
    unsigned audit_max_cell;
    unsigned audit_nonfinite;
} fmdx_thermo_t;

/* two-temperature model description: the reference's 1 x 1 x ncz turi, and
   (no counterpart in the reference, whose 3-D updater is missing,
   ttm.c:451-452) ncx x ncy x ncz turies of type 1, periodic in x and y, with
   constant conductivity or kappa = K * Te / Ti.  Arrays are indexed
   (ix * ncy + iy) * (ncz + 2) + iz with a ghost layer in z only. */
typedef struct
{
    int type;                /* 1: constant K, G, Ce = gamma Te; 2: tabulated Ce(Te), G(Te), Zhigilei K */
    int ncz;                 /* turi cells along z */
    int ratio;               /* TTM sub-steps per MD step */
    double tcellh_z;         /* turi cell height */
    double tcell_volume;
    double C_gamma, K, G;                /* type 1, internal units */
    double Kzh_A, Kzh_B, Kzh_v2;         /* type 2, internal units */
    int nC, nG;                          /* type 2 table lengths */
    const double *T_C, *Ct, *C_DD;       /* type 2 tables and natural-spline second derivatives */
    const double *T_G, *Gt, *G_DD;
    double laser_factor, laser_absdepth_inv, laser_2sig2_inv, laser_t0, laser_tdiff;
    double frontsurf;
    unsigned min_atoms;
    int ncx, ncy;                        /* turi cells along x and y (0 means 1) */
    double tcellh_x, tcellh_y;
    int kmode;                           /* type 1: 0 constant K, 1 kappa = K * Te / Ti */
    int solver3d;                        /* 1: grid solver even for 1 x 1 x ncz (reduction tests) */
} fmdx_ttm_t;

/* devices */
int fmdx_device_count(int *n);
const char *fmdx_strerror(int code);

/* one system: box, cell grid (nc cells of size cellh per dimension) and kinds */
int fmdx_system_create(fmdx_system_t **out, const double l[3], const int pbc[3], const int nc[3],
                       const double cellh[3], int nkinds, const double *mass, int precision);
/* refuses (FMDX_ESTATE) while the system is a member of a batch */
int fmdx_system_destroy(fmdx_system_t *s);
int fmdx_set_precision(fmdx_system_t *s, int precision);

/* pair table: LJ p0 = 4 eps, p1 = sigma^2; Morse p0 = D0, p1 = alpha, p2 = r0; EAM p0 = iloc, p1 = jloc */
int fmdx_set_pair(fmdx_system_t *s, int k1, int k2, int cat, double rc2, double p0, double p1, double p2);
/* EAM tables on the reference's uniform grids: F/F_dd [nelem][nrho] on rho = i*drho,
   rho/rho_dd [nelem][nr2] and phi/phi_dd [nelem][nelem][nr2] on r^2 = i*dr2 */
int fmdx_set_eam(fmdx_system_t *s, int nelem, int nrho, double drho, int nr2, double dr2, double rc2,
                 const double *F, const double *F_dd, const double *rho, const double *rho_dd,
                 const double *phi, const double *phi_dd, const int *elem_of_kind);

/* atoms in cell order: cell_count[c] atoms of interior cell c (LOOP3D order, z fastest),
   x and v interleaved [3*natoms] */
int fmdx_upload(fmdx_system_t *s, unsigned natoms, const unsigned *cell_count, const double *x,
                const double *v, const int *group, const unsigned *atomid, const unsigned *kind);
/* synchronises; any output pointer may be NULL */
int fmdx_download(fmdx_system_t *s, unsigned *natoms, unsigned *cell_count, double *x, double *v,
                  double *f, double *aux, int *group, unsigned *atomid, unsigned *kind);
int fmdx_natoms(fmdx_system_t *s, unsigned *n);   /* synchronises */
int fmdx_ncells(fmdx_system_t *s, unsigned *n);

/* step seams (asynchronous) */
int fmdx_update_forces(fmdx_system_t *s, int active_group, int ttm_friction);
/* the same with flags: FMDX_FORCES_ENERGY also sums the potential energy into the thermo record
   (fmdx_update_forces always does); without it the pair passes skip the pair-energy terms and the
   energy reductions, and the record keeps the energy of the last call that summed it -- for time
   loops whose step is not observed (integrators.c:44-50 needs only the temperature) */
#define FMDX_FORCES_ENERGY 1
int fmdx_update_forces_ex(fmdx_system_t *s, int active_group, int ttm_friction, int flags);
/* fmdx_update_forces without any device allocation: no counterpart in the reference.  For callers
   that cannot allocate (HIP graph capture, serving loops that forbid hipMalloc).  The call's
   scratch -- per-cell energies, the density pass's tile-round records, energy partials -- lives in
   one fixed 16 MiB device array per GPU instead of the system's buffers; the density pass keeps no
   hit words, so the force pass rescans (same results, bit for bit).  Systems of more than
   FMDX_NOALLOC_MAX_CELLS cells are refused with FMDX_ESIZE.  Every call on one GPU shares that
   array: calls that overlap on one GPU (two streams, two threads) race. */
#define FMDX_NOALLOC_MAX_CELLS 590000
int fmdx_update_forces_noalloc(fmdx_system_t *s, int active_group, int ttm_friction);
int fmdx_vv_start(fmdx_system_t *s, int active_group, double dt, int thermostat, double tau, double T0);
int fmdx_vv_finish(fmdx_system_t *s, int active_group, double dt);
int fmdx_se_step(fmdx_system_t *s, int active_group, double dt);
int fmdx_refresh(fmdx_system_t *s);
int fmdx_thermo(fmdx_system_t *s, int active_group);
/* velocity edits on resident atoms (fmd_matt_addVelocity, matter.c:117-150) */
int fmdx_add_velocity(fmdx_system_t *s, int group, double vx, double vy, double vz);

/* synchronises; returns the thermo record and the sticky error word */
int fmdx_read_thermo(fmdx_system_t *s, fmdx_thermo_t *out);
int fmdx_clear_errors(fmdx_system_t *s);
int fmdx_sync(fmdx_system_t *s);
/* state check (asynchronous): counts the atoms over the cells, the fullest cell and the atoms with a
   non-finite position, velocity or force; fmdx_read_thermo returns them (audit_*) */
int fmdx_audit(fmdx_system_t *s);
/* set the per-cell slot capacity (synchronises); the library's own choices are multiples of 32 */
int fmdx_set_capacity(fmdx_system_t *s, unsigned cap);

/* 1-D TTM on the device */
int fmdx_ttm_setup(fmdx_system_t *s, const fmdx_ttm_t *cfg, const double *Te_init /* ncz+2 */);
int fmdx_ttm_set_frontsurf(fmdx_system_t *s, double frontsurf, unsigned min_atoms);
/* atom -> turi-cell sums, presolve and the explicit sub-steps; time = MD time of the update */
int fmdx_ttm_update(fmdx_system_t *s, double time, double dt);
/* the three phases of fmdx_ttm_update, for callers that raise field events in between:
   atom -> turi-cell sums (turi.c:1112-1261), presolve (ttm.c:194-248), sub-steps (ttm.c:250-383) */
int fmdx_ttm_fields(fmdx_system_t *s);
int fmdx_ttm_presolve(fmdx_system_t *s);
int fmdx_ttm_solve(fmdx_system_t *s, double time, double dt);
/* synchronises; arrays of ncx*ncy*(ncz+2) entries (z ghosts included), any may be NULL */
int fmdx_ttm_download(fmdx_system_t *s, double *Te, double *Ti, double *xi, double *vcm /* 3*(ncz+2) */,
                      unsigned *num, double *mass);
int fmdx_ttm_upload_te(fmdx_system_t *s, const double *Te);

/* bounding box of the atoms of a group (fmd_matt_findLimits, matter.c:152-213); synchronises */
int fmdx_find_limits(fmdx_system_t *s, int group, double lo[3], double hi[3]);

/* device bytes held by the system (admission control of served jobs) */
int fmdx_memory(fmdx_system_t *s, size_t *bytes);

/* timing helpers for benchmarks: HIP events on the system stream */
int fmdx_event_record(fmdx_system_t *s, int slot);
int fmdx_event_elapsed_ms(fmdx_system_t *s, int slot0, int slot1, float *ms);
/* per-kernel timing of the last fmdx_update_forces / step (ms), collected when enabled */
int fmdx_profile_enable(fmdx_system_t *s, int on);
int fmdx_profile_read(fmdx_system_t *s, double *ms_by_phase /* 8 entries */, unsigned *count);

/* ------------------------------------------------------------------ batches
 * No counterpart in the reference (one system per fmd_t): independent systems
 * advanced together, every kernel of a step covering all members through
 * (system, block) work lists, the step replayed from a captured HIP graph.
 * For ensembles of small and medium systems and for served jobs.  A member is
 * an uploaded system without TTM, in one batch at a time; while it is a
 * member only the batch calls act on it, and fmdx_system_destroy,
 * fmdx_set_capacity, fmdx_upload and fmdx_set_eam refuse it (FMDX_ESTATE),
 * since the batch's captured step reads its device pointers.  Its trajectory
 * is the same alone and in any mix. */
typedef struct fmdx_batch fmdx_batch_t;

int fmdx_batch_create(fmdx_batch_t **out, int precision);
void fmdx_batch_destroy(fmdx_batch_t *b);
/* join: brings the member's forces and thermo up to date first (integrators.c:232);
   dt, Berendsen flag, tau and T0 as fmd_dync_integrate / fmd_dync_equilibrate take them */
int fmdx_batch_add(fmdx_batch_t *b, fmdx_system_t *s, double dt, int thermostat, double tau, double T0);
int fmdx_batch_remove(fmdx_batch_t *b, fmdx_system_t *s);
int fmdx_batch_size(fmdx_batch_t *b, int *n);
int fmdx_batch_use_graph(fmdx_batch_t *b, int on);
/* status words are read (and slot capacities grown) every `steps` steps; default 8 */
int fmdx_batch_set_check_interval(fmdx_batch_t *b, int steps);
/* Which cells a step visits.  FMDX_COMPACT_AUTO (default): when under 60 % of the members'
   cells hold atoms (vacuum around clusters), the step visits only the occupied cells and
   slot groups, from lists the device rebuilds after every re-binning; fuller batches visit
   every cell.  FMDX_COMPACT_OFF: every cell.  FMDX_COMPACT_TIGHT: always the occupied
   cells, with launch grids exactly as long as the lists were at the last status read
   (cells occupied since then are taken by a looping launch).  Results are the same bits
   in every mode. */
#define FMDX_COMPACT_OFF 0
#define FMDX_COMPACT_AUTO 1
#define FMDX_COMPACT_TIGHT 2
int fmdx_batch_set_compaction(fmdx_batch_t *b, int mode);
/* velocity-Verlet steps of every member; synchronises at its end */
int fmdx_batch_step(fmdx_batch_t *b, int nsteps);
/* thermo record and status of every member, in join order; synchronises */
int fmdx_batch_status(fmdx_batch_t *b, fmdx_thermo_t *out);
int fmdx_batch_sync(fmdx_batch_t *b);

#ifdef __cplusplus
}
#endif

#endif /* FMDX_H */
