/*
 * fmd_oracle.h -- CPU restatement of the reference FMD per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU path
 * is compared against; it is linked or loaded only by tests/, by the smoke
 * check of __graft_entry__ and by the CPU-baseline leg of bench.py.  The
 * product (libfmd.so / libfmdx.so) never calls it.
 *
 * It restates, single rank and single thread (the reference default
 * numthreads=1, misc.c:515), the algorithm of:
 *   grid geometry            misc.c:41-62, potential.c:251-291
 *   cell storage + lists     subdomain.c:76-305, cell.c:98-160, cell.h:84-95
 *   binning / migration      misc.c:242-259, misc.c:268-308, md-ghost.c:41-167,827-869
 *   ghost halo               md-ghost.c:85-107,711-825
 *   LJ / Morse / EAM         lj.c:28-90, lj.h:26-58, morse.c:30-92, morse.h:26-58,
 *                            eam.c:29-322, eam.h:29-406, forces.c:33-393
 *   EAM table preparation    eam.c:324-555, cspline.c:24-117, cspline.h:26-44
 *   integrators              integrators.c:40-264, matter.c:31-68
 *   timers                   timer.c:27-84
 *   1-D TTM                  turi.c:959-1373, ttm.c:63-383, forces.c:292-330
 * Each function below names the lines it follows.
 */
#ifndef FMD_ORACLE_H
#define FMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* units, general.h:37-48 */
#define ORC_MASS_UNIT 9.6485332907e+03
#define ORC_KB 8.6173303e-5

/* potential categories, potential.h:26-32 */
enum { ORC_POT_NONE = 0, ORC_POT_LJ = 1, ORC_POT_EAM = 2, ORC_POT_MORSE = 3 };

orc_t *orc_create(const double l[3], const int pbc[3], int nkinds, const double *mass_amu,
                  const char *const *names);
void orc_free(orc_t *o);

/* potentials */
void orc_set_lj(orc_t *o, int k1, int k2, double sigma, double eps, double cutoff);
void orc_set_morse(orc_t *o, int k1, int k2, double D0, double alpha, double r0, double cutoff);
int orc_load_eam(orc_t *o, const char *path);        /* returns 0 on success */
int orc_apply_eam(orc_t *o, int k1, int k2);         /* uses element names */
double orc_eam_cutoff(orc_t *o);

/* EAM table inspection (for table-level parity tests) */
int orc_eam_sizes(orc_t *o, int *nelem, int *nrho, int *nr2, double *drho, double *dr2, double *rc2);
void orc_eam_tables(orc_t *o, int elem, double *F, double *FDD, double *rho, double *rhoDD);
void orc_eam_phi(orc_t *o, int e1, int e2, double *phi, double *phiDD);

/* grid: must be called after potentials are applied (misc.c:41-62) */
int orc_make_grid(orc_t *o);
void orc_grid_info(orc_t *o, int nc[3], double cellh[3]);

/* atoms are appended in the given order, binned by floor(x/cellh) (misc.c:242-259) */
int orc_add_atoms(orc_t *o, int n, const double *x, const double *v, const int *group, const int *kind);
/* after all atoms are added: global grid -> local grid (matter.c:329-363) */
void orc_distribute(orc_t *o);

int orc_natoms(orc_t *o);
/* atoms in cell storage order: interior cells in LOOP3D order, then within-cell order */
void orc_get_atoms(orc_t *o, double *x, double *v, double *f, double *aux, int *group,
                   unsigned *atomid, int *kind, int *cellidx);
void orc_cell_counts(orc_t *o, int *counts);

/* step seams */
int orc_refresh(orc_t *o);                              /* misc.c:268-308 */
void orc_forces(orc_t *o, int active_group);            /* forces.c:332-382 */
void orc_vv_start(orc_t *o, int active_group, double dt, int thermostat, double tau, double T0);
void orc_vv_finish(orc_t *o, int active_group, double dt);
void orc_se_step(orc_t *o, int active_group, double dt);
void orc_thermo(orc_t *o, int active_group);            /* matter.c:31-68 */
/* out: [0]=PE [1]=KE [2]=T [3..5]=momentum (internal units) [6]=N [7]=FembSum */
void orc_get_energies(orc_t *o, double *out);
int orc_error_count(orc_t *o, int *last_code);

/* whole loops (integrators.c:181-264); return number of steps taken */
int orc_integrate_vv(orc_t *o, int active_group, double *time, double duration, double dt,
                     int thermostat, double tau, double T0);

/* timer predicate, timer.c:27-36 */
int orc_timer_is_its_time(double t, double dt_half, double start, double interval);

/* TTM.  The reference's 1-D turi (1 x 1 x ncz, type 1 and 2), and this
   build's extension to ncx x ncy x ncz turies periodic in x and y, type 1
   with constant conductivity (kmode 0) or kappa = K * Te / Ti (kmode 1).
   parity unpinned by the reference for the extension itself: it is pinned by
   reduction to the 1-D scheme, by the type-2 special case (Zhigilei A = 0,
   linear Ce table, constant G table) and by the flux-conservation identity. */
typedef struct
{
    int type;              /* 1 or 2 */
    int ncz;               /* global turi cells along z */
    int ratio;             /* timestep ratio */
    double C_gamma, K, G;  /* type 1 (internal units) */
    double Kzh_A, Kzh_B, Kzh_v2;        /* type 2 */
    int nC, nG;
    const double *T_C, *Ct, *T_G, *Gt;  /* type 2 tables (internal units) */
    double act_frac;
    double laser_factor, laser_absdepth_inv, laser_2sig2_inv, laser_t0, laser_tdiff;
    double start, stop;
    int ncx, ncy;          /* turi cells along x and y (0 means 1) */
    int kmode;             /* 0: constant K; 1: K * Te / Ti */
    int solver3d;          /* 1: use the ncx x ncy x ncz solver even for a 1 x 1 x ncz turi */
} orc_ttm_cfg_t;
int orc_ttm_init(orc_t *o, const orc_ttm_cfg_t *cfg, double Te_init);
void orc_ttm_get_ready(orc_t *o, int active_group);
void orc_ttm_update(orc_t *o, double time, double dt);   /* fields + presolve + solve */
/* field injection and the FD sub-steps alone (pin tests of the solver) */
void orc_ttm_set_fields(orc_t *o, const double *Te, const double *Ti, const unsigned *num, double frontsurf,
                        unsigned min_atoms);
void orc_ttm_solve_only(orc_t *o, double time, double dt);
void orc_ttm_presolve_solve(orc_t *o, double time, double dt);
/* per turi-cell arrays, length ncx*ncy*(ncz+2), index (ix*ncy + iy)*(ncz+2) + iz (z ghosts included) */
void orc_ttm_arrays(orc_t *o, double *Te, double *Ti, double *xi, double *vcmz, unsigned *num,
                    double *mass);
void orc_ttm_params(orc_t *o, double *frontsurf, unsigned *min_atoms, double *tcellh_z);
int orc_integrate_se(orc_t *o, int active_group, double *time, double duration, double dt);

/* stand-alone spline helpers (cspline.c:50-103) */
void orc_spline_prepare(double h, const double *y, int n, double *yDD);
void orc_spline_prepare_adv(const double *x, const double *y, int n, double *yDD);
double orc_spline_val(double h, const double *ya, const double *yDDa, double x);
double orc_spline_deriv(double h, const double *ya, const double *yDDa, double x);

#ifdef __cplusplus
}
#endif

#endif /* FMD_ORACLE_H */
