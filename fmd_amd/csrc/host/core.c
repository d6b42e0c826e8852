/*
 * core.c -- instance lifetime, error events, timers, process and I/O
 * settings, version, 3-D output arrays, random numbers.
 *
 * Reference behaviour followed:
 *   fmd_create / fmd_free          misc.c:499-558, 619-636
 *   error events                   error.c:27-236 (message to stderr when
 *                                  enabled, then FMD_EVENT_ERROR; execution
 *                                  continues)
 *   timers                         timer.c:27-84
 *   process / I-O setters          misc.c:424-617, subdomain.c:325-341
 *   version                        version.c:25-61
 * Differences: there is one process and one GPU, so fmd_box_setSubdomains
 * accepts any layout and reports "is an MD process"; random numbers come
 * from a seeded generator (FMD_SEED) instead of time(NULL).
 */
#include <stdarg.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include "fmd_host.h"

#define FORMAT1 "FMD (%s, %s, %d): ERROR: "

static double wall_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------- errors */

static void raise_error(fmd_t *md, fmd_error_t code, bool major, const char *src, const char *func, int line,
                        void *p1, void *p2, void *p3)
{
    if (md->EventHandler == NULL) return;
    fmd_event_params_error_t e;
    e.error = code;
    e.major = major;
    e.source = (fmd_string_t)src;
    e.func = (fmd_string_t)func;
    e.line = line;
    e.p1 = p1;
    e.p2 = p2;
    e.p3 = p3;
    md->EventHandler(md, FMD_EVENT_ERROR, md->userobject, (fmd_params_t *)&e);
}

void fmdh_err_unable_open_file(fmd_t *md, bool major, const char *src, const char *func, int line, const char *path)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "Unable to open the file (%s)!\n", src, func, line, path);
    raise_error(md, FMD_ERR_UNABLE_OPEN_FILE, major, src, func, line, (void *)path, NULL, NULL);
}

void fmdh_err_unexpected_position(fmd_t *md, bool major, const char *src, const char *func, int line)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "Unexpected particle position!\n", src, func, line);
    raise_error(md, FMD_ERR_UNEXPECTED_POSITION, major, src, func, line, NULL, NULL, NULL);
}

void fmdh_err_outside_real_interval(fmd_t *md, bool major, const char *src, const char *func, int line,
                                    const char *qname, double val, const char *intv)
{
    if (md->ShowErrorMessages)
        fprintf(stderr, FORMAT1 "The %s (%g) is outside the interval %s!\n", src, func, line, qname, val, intv);
    raise_error(md, FMD_ERR_OUTSIDE_REAL_INTERVAL, major, src, func, line, (void *)qname, &val, (void *)intv);
}

void fmdh_err_unacceptable_int(fmd_t *md, bool major, const char *src, const char *func, int line,
                               const char *qname, int val)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "Unacceptable %s (%d)!\n", src, func, line, qname, val);
    raise_error(md, FMD_ERR_UNACCEPTABLE_INT_VALUE, major, src, func, line, (void *)qname, &val, NULL);
}

void fmdh_err_file_corrupted(fmd_t *md, bool major, const char *src, const char *func, int line,
                             const char *ftype, const char *path)
{
    if (md->ShowErrorMessages)
        fprintf(stderr, FORMAT1 "Not a healthy %s file (%s)!\n", src, func, line, ftype, path);
    raise_error(md, FMD_ERR_FILE_CORRUPTED, major, src, func, line, (void *)ftype, (void *)path, NULL);
}

void fmdh_err_unable_allocate(fmd_t *md, bool major, const char *src, const char *func, int line, size_t sz)
{
    if (md->ShowErrorMessages)
        fprintf(stderr, FORMAT1 "Unable to allocate memory (%zu bytes)!\n", src, func, line, sz);
    raise_error(md, FMD_ERR_UNABLE_ALLOCATE_MEM, major, src, func, line, &sz, NULL, NULL);
}

void fmdh_err_function_failed(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "Function failed (%s)!\n", src, func, line, what);
    raise_error(md, FMD_ERR_FUNCTION_FAILED, major, src, func, line, (void *)what, NULL, NULL);
}

void fmdh_err_not_supported(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "%s\n", src, func, line, what);
    raise_error(md, FMD_ERR_NOT_SUPPORTED_YET, major, src, func, line, (void *)what, NULL, NULL);
}

void fmdh_err_unprepared(fmd_t *md, bool major, const char *src, const char *func, int line, const char *what)
{
    if (md->ShowErrorMessages) fprintf(stderr, FORMAT1 "The %s is not prepared!\n", src, func, line, what);
    raise_error(md, FMD_ERR_UNPREPARED, major, src, func, line, (void *)what, NULL, NULL);
}

void fmdh_err_wrong_potential(fmd_t *md, bool major, const char *src, const char *func, int line, const char *pot,
                              unsigned k1, unsigned k2)
{
    if (md->ShowErrorMessages)
        fprintf(stderr, FORMAT1 "The %s potential cannot be applied to atom kinds %u and %u!\n", src, func, line,
                pot, k1, k2);
    raise_error(md, FMD_ERR_WRONG_POTENTIAL, major, src, func, line, (void *)pot, &k1, &k2);
}

/* -------------------------------------------------------------- instance */

fmd_t *fmd_create()
{
    fmd_t *md = calloc(1, sizeof(fmd_t));
    if (!md) return NULL;
    md->numthreads = 1;
    md->SaveDirectory = calloc(1, 1);
    md->ActiveGroup = FMD_GROUP_ALL;
    md->SaveConfigMode = FMD_SCM_XYZ_ATOMSNUM;
    md->_OldNumberOfParticles = -1;
    md->ShowErrorMessages = true;
    md->active_ttm_turi = -1;
    md->precision = FMDX_PREC_FP64;
    const char *prec = getenv("FMD_PRECISION");
    if (prec && (strcmp(prec, "mixed") == 0 || strcmp(prec, "1") == 0)) md->precision = FMDX_PREC_MIXED;
    fmd_proc_setCellIncrement(md, 3);
    fmdh_rng_reseed(md);
    md->WallTimeOrigin = wall_now();
    return md;
}

void fmd_free(fmd_t *md)
{
    if (!md) return;
    if (md->dev && fmdx_system_destroy(md->dev) != FMDX_OK)
        fmdh_err_function_failed(md, true, HERE, "fmdx_system_destroy: the system is still in a batch");
    if (md->ConfigFilep && md->_OldNumberOfParticles != -1) fclose(md->ConfigFilep);
    fmdh_gg_free(md);
    fmdh_potsys_free(md);
    fmdh_turi_free_all(md);
    free(md->timers);
    free(md->SaveDirectory);
    free(md);
}

void fmd_setEventHandler(fmd_t *md, void *usp, fmd_EventHandler_t func)
{
    md->EventHandler = func;
    md->userobject = usp;
}

/* ------------------------------------------------------------------ box */

void fmd_box_setPBC(fmd_t *md, bool PBCx, bool PBCy, bool PBCz)
{
    md->PBC[0] = PBCx;
    md->PBC[1] = PBCy;
    md->PBC[2] = PBCz;
    md->PBCdetermined = true;
}

void fmd_box_setSize(fmd_t *md, fmd_real_t sx, fmd_real_t sy, fmd_real_t sz)
{
    if (md->grid_made) return;   /* misc.c:560-569: ignored once the grid exists */
    md->l[0] = sx;
    md->l[1] = sy;
    md->l[2] = sz;
    md->BoxSizeDetermined = true;
}

bool fmd_box_setSubdomains(fmd_t *md, int dimx, int dimy, int dimz)
{
    md->ns[0] = dimx;
    md->ns[1] = dimy;
    md->ns[2] = dimz;
    return true;
}

/* --------------------------------------------------------------- process */

fmd_real_t fmd_proc_getWallTime(fmd_t *md) { return wall_now() - md->WallTimeOrigin; }
bool fmd_proc_hasSubdomain(fmd_t *md) { (void)md; return true; }
bool fmd_proc_isRoot(fmd_t *md) { (void)md; return true; }
void fmd_proc_setNumThreads(fmd_t *md, int num) { md->numthreads = num; }
int fmd_proc_getCellIncrement(fmd_t *md) { return md->cell_incm1 + 1; }

void fmd_proc_setCellIncrement(fmd_t *md, int incr)
{
    if (incr <= 0)
    {
        fmdh_err_unacceptable_int(md, false, HERE, "cell increment value", incr);
        return;
    }
    md->cell_incm1 = incr - 1;
    md->cell_2incm1 = 2 * incr - 1;
}

/* ------------------------------------------------------------------ I/O */

void fmd_io_setSaveDirectory(fmd_t *md, fmd_string_t directory)
{
    char *s = realloc(md->SaveDirectory, strlen(directory) + 1);
    if (!s) return;
    md->SaveDirectory = s;
    strcpy(md->SaveDirectory, directory);
}

void fmd_io_setSaveConfigMode(fmd_t *md, fmd_SaveConfigMode_t mode) { md->SaveConfigMode = mode; }
void fmd_io_setShowErrorMessages(fmd_t *md, bool show) { md->ShowErrorMessages = show; }

void fmd_io_printf(fmd_t *md, const fmd_string_t restrict format, ...)
{
    (void)md;
    va_list ap;
    va_start(ap, format);
    vprintf(format, ap);
    va_end(ap);
}

/* ---------------------------------------------------------------- timers */

bool fmdh_timer_is_its_time(double t, double dt_half, double start, double interval)
{
    /* _fmd_timer_is_its_time, timer.c:27-36 */
    double tc = round((t - start) / interval) * interval + start;
    double d = t - tc;
    return d >= -dt_half && d < dt_half;
}

fmd_handle_t fmd_timer_makeSimple(fmd_t *md, fmd_real_t start, fmd_real_t interval, fmd_real_t stop)
{
    timer_t_ *t = realloc(md->timers, (md->ntimers + 1) * sizeof(timer_t_));
    if (!t) return -1;
    md->timers = t;
    t[md->ntimers].enabled = true;
    t[md->ntimers].start = start;
    t[md->ntimers].interval = interval;
    t[md->ntimers].stop = stop;
    return md->ntimers++;
}

/* whether fmdh_send_timer_ticks would send a tick at time t (the same predicates) */
bool fmdh_tick_due(fmd_t *md, double t)
{
    for (unsigned i = 0; i < md->ntimers; i++)
    {
        timer_t_ *tm = &md->timers[i];
        if (tm->enabled && t >= tm->start - md->timestep / 2.0 &&
            (t < tm->stop + md->timestep / 2.0 || tm->stop < tm->start) &&
            fmdh_timer_is_its_time(t, md->timestep / 2.0, tm->start, tm->interval))
            return true;
    }
    return false;
}

/* _fmd_timer_sendTimerTickEvents, timer.c:64-84 */
void fmdh_send_timer_ticks(fmd_t *md)
{
    for (unsigned i = 0; i < md->ntimers; i++)
    {
        timer_t_ *t = &md->timers[i];
        if (t->enabled && md->time >= t->start - md->timestep / 2.0 &&
            (md->time < t->stop + md->timestep / 2.0 || t->stop < t->start))
        {
            if (fmdh_timer_is_its_time(md->time, md->timestep / 2.0, t->start, t->interval))
            {
                fmd_event_params_timer_tick_t p;
                p.timer = (fmd_handle_t)i;
                md->EventHandler(md, FMD_EVENT_TIMER_TICK, md->userobject, (fmd_params_t *)&p);
            }
        }
    }
}

/* --------------------------------------------------------------- version */

int fmd_version_getMajor() { return 0; }
int fmd_version_getMinor() { return 4; }
int fmd_version_getRevision() { return 0; }
char fmd_version_getType() { return 'r'; }

fmd_string_t fmd_version_getString()
{
    char *out = malloc(32);
    if (out) snprintf(out, 32, "%d.%d.%d", fmd_version_getMajor(), fmd_version_getMinor(), fmd_version_getRevision());
    return out;
}

/* ------------------------------------------------------------ 3-D arrays */

/* one allocation: data[d0] -> rows[d0*d1] -> elements, indexable as a[i][j][k] */
fmd_array3s_t *fmdh_array3s_create(unsigned d0, unsigned d1, unsigned d2, size_t elsize)
{
    fmd_array3s_t *a = malloc(sizeof *a);
    if (!a) return NULL;
    size_t np0 = d0 * sizeof(void **), np1 = (size_t)d0 * d1 * sizeof(void *);
    size_t head = (np0 + np1 + 15) & ~(size_t)15;
    char *blk = calloc(1, head + (size_t)d0 * d1 * d2 * elsize + 16);
    if (!blk) { free(a); return NULL; }
    void ***p0 = (void ***)blk;
    void **p1 = (void **)(blk + np0);
    char *data = blk + head;
    for (unsigned i = 0; i < d0; i++)
    {
        p0[i] = p1 + (size_t)i * d1;
        for (unsigned j = 0; j < d1; j++) p1[(size_t)i * d1 + j] = data + ((size_t)i * d1 + j) * d2 * elsize;
    }
    a->data = p0;
    a->dims[0] = d0;
    a->dims[1] = d1;
    a->dims[2] = d2;
    return a;
}

void fmd_array3s_free(fmd_array3s_t *array)
{
    if (array)
    {
        free(array->data);
        free(array);
    }
}

/* --------------------------------------------------------- random numbers */

/* splitmix64 stream; gaussian by the Marsaglia polar method */
static uint64_t rng_next(fmd_t *md)
{
    uint64_t z = (md->rng_seed += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    md->rng_calls++;
    return z ^ (z >> 31);
}

void fmdh_rng_reseed(fmd_t *md)
{
    const char *s = getenv("FMD_SEED");
    md->rng_seed = s ? strtoull(s, NULL, 10) : 5489ull;
}

double fmdh_rng_uniform(fmd_t *md)
{
    return (rng_next(md) >> 11) * (1.0 / 9007199254740992.0);
}

double fmdh_rng_gaussian(fmd_t *md, double sigma)
{
    double u, v, s;
    do
    {
        u = 2.0 * fmdh_rng_uniform(md) - 1.0;
        v = 2.0 * fmdh_rng_uniform(md) - 1.0;
        s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    return sigma * u * sqrt(-2.0 * log(s) / s);
}
