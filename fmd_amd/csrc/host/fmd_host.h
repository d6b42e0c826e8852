/*
 * fmd_host.h -- internal state of the host side of libfmd.so.
 *
 * The host keeps the reference's configuration state (box, kinds, pair
 * table, timers, turies, I/O settings) and, until the first dynamics call,
 * a host-side "global grid" of cells filled in generation / file order
 * (misc.c:41-62, 242-259; structure.c:95-129).  The first dynamics call
 * uploads that grid to the device (the single-process counterpart of
 * _fmd_matt_distribute, matter.c:399-447); from then on the atoms live in
 * HBM and every per-step operation is a call into libfmdx.so.
 */
#ifndef FMD_HOST_H
#define FMD_HOST_H

#include <stdbool.h>
#include <stdio.h>
#include "fmd.h"
#include "fmdx.h"

/* units and constants, reference general.h:37-48 */
#define MD_MASS_UNIT 9.6485332907e+03
#define METER 1e10
#define SECOND 1e12
#define METER_PER_SECOND 1e-2
#define JOULE_PER_METER3_KELVIN 6.2415091259e-12
#define JOULE_PER_METER3_KELVIN2 6.2415091259e-12
#define WATT_PER_METER_KELVIN 6.2415091259e-4
#define WATT_PER_METER2 6.2415091259e-14
#define WATT_PER_METER3_KELVIN 6.2415091259e-24
#define K_BOLTZMANN 8.6173303e-5

#define MAX_PATH_LENGTH 256
#define HMAXK FMDX_MAX_KINDS
#define HMAXE FMDX_MAX_ELEMENTS

enum { POT_NONE = 0, POT_LJ_6_12 = 1, POT_EAM_ALLOY = 2, POT_MORSE = 3 };

struct _fmd_pot
{
    int cat;
    void *data;
};

typedef struct
{
    double eps, eps4, sig, sig_sqr, cutoff_sqr;
} lj_t;

typedef struct
{
    double D0, alpha, r0, cutoff_sqr;
} morse_t;

typedef struct
{
    char name[64];
    double mass, lattice_a;
    double *F, *Fdd;           /* [Nrho] */
    double *rho, *rhodd;        /* [Nr2] on the r^2 grid */
} eam_element_t;

typedef struct
{
    int nelem, Nrho, Nr, Nr2;
    double drho, dr, dr2, cutoff_sqr;
    eam_element_t el[HMAXE];
    double *phi[HMAXE][HMAXE], *phidd[HMAXE][HMAXE];
} eam_t;

typedef struct
{
    int cat;
    void *data;
    int iloc, jloc;
} potpair_t;

typedef struct
{
    double mass;         /* internal units */
    char *name;
    int eam_element;     /* -1 when the kind has no EAM element */
} atomkind_t;

/* host-side cell of the global grid (before the atoms move to the device) */
typedef struct
{
    unsigned n, cap;
    double *x, *v;
    int *grp;
    unsigned *aid, *kind;
} hcell_t;

typedef struct
{
    bool enabled;
    double start, stop, interval;
} timer_t_;

/* fields of a turi, reference turi.h:33-52 */
typedef struct
{
    int cat;
    int timestep;
    double time;
    int ndep, dep[2];
    int nint, nint_all;
    double intervals[16], intervals_all[16];
    bool allhave_now;
} field_t_;

typedef struct
{
    int cat;
    unsigned tdims[3];
    double tcellh[3], tcell_volume;
    double starttime, stoptime;
    int nfields;
    field_t_ fields[16];
    /* TTM */
    bool is_ttm;
    int ttm_type, ratio;
    double C_gamma, K, G, Kzh_A, Kzh_B, Kzh_v2;
    int nC, nG;
    double *T_C, *Ct, *C_DD, *T_G, *Gt, *G_DD;
    double CellActivFrac;
    double laser_factor, laser_absdepth_inv, laser_2sig2_inv, laser_t0, laser_tdiff;
    double frontsurf;
    unsigned initial_atoms, min_atoms;
    int kmode;                   /* type 1: 0 constant K, 1 K * Te / Ti */
    bool grid_solver;            /* grid solver forced for a 1 x 1 x n turi */
    unsigned ncells;             /* tdims[0] * tdims[1] * (tdims[2] + 2) */
    double *Te_host;             /* Te set before the device copy exists [ncells] */
    bool dev_ready;
    int iTe, ixi, inum, ivcm, iTi, imass;
} turi_t_;

struct _fmd_array3s
{
    void ***data;
    fmd_utriple_t dims;
};

struct _fmd
{
    /* box */
    bool BoxSizeDetermined, PBCdetermined;
    double l[3];
    int PBC[3];
    int ns[3];
    int nc[3];
    double cellh[3];
    bool grid_made;
    /* kinds and potentials */
    unsigned nkinds;
    atomkind_t *kinds;
    potpair_t *pottable;          /* [nkinds*nkinds] */
    struct _fmd_pot **potlist;
    int npots;
    bool potcats_valid;
    int potcats_num;
    bool eam_applied;
    eam_t *eam;                   /* the EAM potential applied (at most one) */
    /* atoms before distribution */
    hcell_t *gg;
    unsigned ngg;
    unsigned TotalNoOfParticles;
    bool ParticlesDistributed;
    /* device */
    fmdx_system_t *dev;
    int precision;
    unsigned dev_err_reported;
    bool audit_reported;          /* a failed device state check has been reported */
    bool dev_pot_dirty;
    /* dynamics */
    double time, timestep;
    int time_iteration;
    int ActiveGroup;
    double DesiredTemperature, BerendsenThermostatParam;
    bool KineticEnergyUpdated;
    bool thermo_host_valid;
    double GroupTemperature, GroupKineticEnergy, GroupPotentialEnergy;
    double GroupMomentum[3];
    unsigned GroupParticlesNum;
    /* events */
    fmd_EventHandler_t EventHandler;
    void *userobject;
    timer_t_ *timers;
    unsigned ntimers;
    turi_t_ *turies;
    unsigned nturies;
    int active_ttm_turi;          /* -1 when none */
    /* I/O */
    char *SaveDirectory;
    fmd_SaveConfigMode_t SaveConfigMode;
    FILE *ConfigFilep;
    int _OldNumberOfParticles;
    int _FileIndex;
    bool ShowErrorMessages;
    int cell_incm1, cell_2incm1;
    int numthreads;
    double WallTimeOrigin;
    unsigned long long rng_seed;
    unsigned long long rng_calls;
};

/* errors (reference error.c) */
void fmdh_err_unable_open_file(fmd_t *md, bool major, const char *src, const char *func, int line, const char *path);
void fmdh_err_unexpected_position(fmd_t *md, bool major, const char *src, const char *func, int line);
void fmdh_err_outside_real_interval(fmd_t *md, bool major, const char *src, const char *func, int line,
                                    const char *qname, double val, const char *intv);
void fmdh_err_unacceptable_int(fmd_t *md, bool major, const char *src, const char *func, int line,
                               const char *qname, int val);
void fmdh_err_file_corrupted(fmd_t *md, bool major, const char *src, const char *func, int line,
                             const char *ftype, const char *path);
void fmdh_err_unable_allocate(fmd_t *md, bool major, const char *src, const char *func, int line, size_t sz);
void fmdh_err_function_failed(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what);
void fmdh_err_not_supported(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what);
void fmdh_err_unprepared(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what);
void fmdh_err_wrong_potential(fmd_t *md, bool major, const char *src, const char *func, int line, const char *pot,
                              unsigned k1, unsigned k2);

#define HERE __FILE__, __func__, __LINE__

/* cubic splines (reference cspline.c) */
void fmdh_spline_prepare(double h, const double *y, int n, double *yDD);
void fmdh_spline_prepare_adv(const double *x, const double *y, int n, double *yDD);
double fmdh_spline_val(double h, const double *ya, const double *yDDa, double x);

/* potentials */
double fmdh_largest_cutoff(fmd_t *md);
bool fmdh_update_potcats(fmd_t *md);
void fmdh_potsys_free(fmd_t *md);

/* grid and host cells */
bool fmdh_make_grid(fmd_t *md);
unsigned fmdh_cell_new(fmd_t *md, hcell_t *c);
void fmdh_cell_remove(hcell_t *c, unsigned i);
void fmdh_gg_free(fmd_t *md);

/* device */
bool fmdh_distribute(fmd_t *md);
bool fmdh_push_potentials(fmd_t *md);
void fmdh_sync_thermo(fmd_t *md);
void fmdh_check_device(fmd_t *md);
bool fmdh_download_all(fmd_t *md, unsigned *n, double **x, double **v, int **grp, unsigned **aid, unsigned **kind,
                       unsigned **cellcnt);

/* dynamics */
void fmdh_send_timer_ticks(fmd_t *md);
bool fmdh_tick_due(fmd_t *md, double t);
bool fmdh_timer_is_its_time(double t, double dt_half, double start, double interval);

/* turies */
void fmdh_turies_update(fmd_t *md, bool X, bool V, bool F);
void fmdh_ttm_get_ready(fmd_t *md);
bool fmdh_ttm_friction_active(fmd_t *md);
void fmdh_turi_free_all(fmd_t *md);

/* random numbers (seeded; FMD_SEED in the environment overrides the seed) */
double fmdh_rng_uniform(fmd_t *md);
double fmdh_rng_gaussian(fmd_t *md, double sigma);
void fmdh_rng_reseed(fmd_t *md);

/* 3-D arrays for field output */
fmd_array3s_t *fmdh_array3s_create(unsigned d0, unsigned d1, unsigned d2, size_t elsize);

#endif /* FMD_HOST_H */
