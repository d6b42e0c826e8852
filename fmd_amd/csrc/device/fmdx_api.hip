/*
 * fmdx_api.hip -- the C-ABI entry points of libfmdx.so (declared in fmdx.h).
 *
 * Owns the device memory of one system, converts the reference's spline
 * tables into per-interval coefficients, and enqueues the step kernels on
 * the system's stream.  Nothing here computes physics on the host.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>
#include "fmdx_sys.h"

/* Streams of destroyed systems are kept for the next system created on the same device:
   creating and destroying a HIP stream cost about 0.2 and 0.45 ms, which a server that
   builds and retires a system per job pays every job.  A stream enters the pool idle. */
static std::mutex g_stream_mu;
static std::vector<std::pair<int, hipStream_t>> g_streams;

static hipError_t stream_take(int dev, hipStream_t *st)
{
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (size_t i = g_streams.size(); i-- > 0;)
            if (g_streams[i].first == dev)
            {
                *st = g_streams[i].second;
                g_streams.erase(g_streams.begin() + i);
                return hipSuccess;
            }
    }
    return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

static void stream_give(int dev, hipStream_t st)
{
    if (hipStreamSynchronize(st) != hipSuccess)
    {
        hipStreamDestroy(st);
        return;
    }
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_streams.emplace_back(dev, st);
}

extern "C" const char *fmdx_strerror(int code)
{
    switch (code)
    {
        case FMDX_OK: return "ok";
        case FMDX_EHIP: return "HIP runtime call failed";
        case FMDX_EARG: return "invalid argument";
        case FMDX_ENOMEM: return "device allocation failed";
        case FMDX_ESTATE: return "call out of order";
        case FMDX_ENODEV: return "no GPU visible";
        case FMDX_ESIZE: return "system larger than the fixed scratch of the allocation-free entry point";
    }
    return "unknown";
}

extern "C" int fmdx_device_count(int *n)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return FMDX_OK;
}

/* per-slot arrays of a layout with capacity cap (ns = ncell * cap slots), mover staging
   for nmov atoms and the per-block thermo partials, carved from D.slot_mem */
static void slot_layout(KDesc &D, Carve &c, size_t ns, size_t nmov, size_t npart)
{
    c(&D.x, ns); c(&D.y, ns); c(&D.z, ns);
    c(&D.vx, ns); c(&D.vy, ns); c(&D.vz, ns);
    c(&D.fx, ns); c(&D.fy, ns); c(&D.fz, ns);
    c(&D.aux, ns);
    c(&D.xq, ns);
    c(&D.grp, ns); c(&D.aid, ns); c(&D.kind, ns);
    c(&D.mv, ns); c(&D.code, ns);
    c(&D.bkey, ns); c(&D.bid, ns);
    c(&D.mvx, nmov); c(&D.mvy, nmov); c(&D.mvz, nmov);
    c(&D.mvvx, nmov); c(&D.mvvy, nmov); c(&D.mvvz, nmov);
    c(&D.mvgrp, nmov); c(&D.mvaid, nmov); c(&D.mvkind, nmov);
    c(&D.part, npart);
}

static void free_slots(KDesc &D)
{
    dfree(D.np);
    carve_free(&D.slot_mem, [&](Carve &c) { slot_layout(D, c, 0, 0, 0); });
}

static int alloc_slots(KDesc &D, int cap, unsigned nmov)
{
    const size_t ns = (size_t)D.ncell * cap;
    const size_t nblk = ((size_t)D.ncell + CPB - 1) / CPB + 1;
    D.cap = cap;
    return carve_alloc(&D.slot_mem, [&](Carve &c) { slot_layout(D, c, ns, nmov, nblk * 8); });
}

/* per-cell arrays, counters, reduction partials and the pair table, carved from s->cell_mem */
static void cell_layout(fmdx_system_t *s, Carve &c, size_t ncell)
{
    KDesc &D = s->D;
    c(&D.cnt, ncell);
    c(&D.nleave, ncell); c(&D.narr, ncell);
    c(&D.nearly, ncell); c(&D.fill, ncell);
    c(&D.counters, NCOUNTERS);
    c(&D.cell_pe, ncell); c(&D.cell_emb, ncell);
    c(&D.htrips, ncell);
    c(&D.htile, ncell);
    c(&D.btouch, ncell / TPB + 1);
    c(&D.epart, 2 * NPART_MAX);
    c(&D.thermo, 16);
    c(&s->dpairs, MAXK * MAXK);
}

/* the EAM tables, carved from s->tab_mem: nF knots of F, nR of rho, nP of phi (each in
   FP64 and FP32), nRF of the r^2 tables and nK of the FP32 derivative knots */
struct TabSizes
{
    size_t nF, nR, nP, nRF, nK;
};

static void tab_layout(fmdx_system_t *s, Carve &c, const TabSizes &z)
{
    c(&s->rf, z.nRF); c(&s->rf32, z.nRF); c(&s->kn32, z.nK);
    c(&s->Fc, z.nF); c(&s->rhoc, z.nR); c(&s->phic, z.nP);
    c(&s->Fc32, z.nF); c(&s->rhoc32, z.nR); c(&s->phic32, z.nP);
}

extern "C" int fmdx_system_create(fmdx_system_t **out, const double l[3], const int pbc[3], const int nc[3],
                                  const double cellh[3], int nkinds, const double *mass, int precision)
{
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return FMDX_ENODEV;
    if (nkinds < 1 || nkinds > MAXK) return FMDX_EARG;
    for (int d = 0; d < 3; d++)
        if (nc[d] < 1 || !(cellh[d] > 0) || !(l[d] > 0)) return FMDX_EARG;
    long ncell = (long)nc[0] * nc[1] * nc[2];
    if (ncell > (1L << 30)) return FMDX_EARG;

    fmdx_system_t *s = (fmdx_system_t *)calloc(1, sizeof(fmdx_system_t));
    if (!s) return FMDX_ENOMEM;
    CK(hipGetDevice(&s->dev));
    if (stream_take(s->dev, &s->st) != hipSuccess) { free(s); return FMDX_EHIP; }
    s->precision = precision;
    s->D.xq_keep = precision == FMDX_PREC_MIXED;
    s->xq_fresh = 0;
    KDesc &D = s->D;
    D.nc0 = nc[0]; D.nc1 = nc[1]; D.nc2 = nc[2];
    D.ncell = (int)ncell;
    D.pbc0 = pbc[0]; D.pbc1 = pbc[1]; D.pbc2 = pbc[2];
    D.L0 = l[0]; D.L1 = l[1]; D.L2 = l[2];
    D.h0 = cellh[0]; D.h1 = cellh[1]; D.h2 = cellh[2];
    D.nkinds = nkinds;
    D.single_cat = -1;
    D.energy = 1;
    for (int k = 0; k < MAXK; k++)
    {
        D.mass[k] = k < nkinds ? mass[k] : 1.0;
        s->mass[k] = D.mass[k];
        D.elem[k] = -1;
    }
    if (carve_alloc(&s->cell_mem, [&](Carve &c) { cell_layout(s, c, ncell); })) { fmdx_system_destroy(s); return FMDX_ENOMEM; }
    hipMemsetAsync(D.cnt, 0, sizeof(int) * ncell, s->st);
    hipMemsetAsync(D.nleave, 0, sizeof(int) * ncell, s->st);
    hipMemsetAsync(D.narr, 0, sizeof(int) * ncell, s->st);
    hipMemsetAsync(D.nearly, 0, sizeof(int) * ncell, s->st);
    hipMemsetAsync(D.fill, 0, sizeof(int) * ncell, s->st);
    hipMemsetAsync(D.counters, 0, sizeof(unsigned) * NCOUNTERS, s->st);
    hipMemsetAsync(D.cell_pe, 0, sizeof(double) * ncell, s->st);
    hipMemsetAsync(D.cell_emb, 0, sizeof(double) * ncell, s->st);
    hipMemsetAsync(D.htrips, 0xff, sizeof(int) * ncell, s->st);   /* -1: no recorded hits */
    hipMemsetAsync(D.htile, 0, sizeof(unsigned long long) * ncell, s->st);
    hipMemsetAsync(D.btouch, 0, sizeof(int) * (ncell / TPB + 1), s->st);
    hipMemsetAsync(D.thermo, 0, sizeof(double) * 16, s->st);
    memset(s->hp, 0, sizeof s->hp);
    D.pairs = s->dpairs;
    for (int i = 0; i < 16; i++) hipEventCreate(&s->ev[i]);
    s->nev = 16;
    for (int i = 0; i < NPROF * 3; i++) hipEventCreate(&s->pev[i]);
    if (hipStreamSynchronize(s->st) != hipSuccess) { fmdx_system_destroy(s); return FMDX_EHIP; }
    *out = s;
    return FMDX_OK;
}

extern "C" int fmdx_system_destroy(fmdx_system_t *s)
{
    if (!s) return FMDX_OK;
    if (s->member) return FMDX_ESTATE;   /* a batch still holds its device pointers */
    if (s->st) hipStreamSynchronize(s->st);
    KDesc &D = s->D;
    free_slots(D);
    carve_free(&s->cell_mem, [&](Carve &c) { cell_layout(s, c, 0); });
    dfree(D.hbits);
    carve_free(&s->tab_mem, [&](Carve &c) { tab_layout(s, c, TabSizes{}); });
    for (int i = 0; i < s->nttm_mem; i++) hipFree(s->ttm_mem[i]);
    for (int i = 0; i < s->nev; i++) hipEventDestroy(s->ev[i]);
    for (int i = 0; i < NPROF * 3; i++)
        if (s->pev[i]) hipEventDestroy(s->pev[i]);
    if (s->st) stream_give(s->dev, s->st);
    free(s);
    return FMDX_OK;
}

extern "C" int fmdx_set_precision(fmdx_system_t *s, int precision)
{
    if (!s || (precision != FMDX_PREC_FP64 && precision != FMDX_PREC_MIXED)) return FMDX_EARG;
    if (precision == FMDX_PREC_MIXED && s->D.has_eam && !s->rhoc32) return FMDX_ESTATE;
    s->precision = precision;
    s->D.xq_keep = precision == FMDX_PREC_MIXED;
    s->xq_fresh = 0;
    return ensure_n3(s, precision);
}

static void refresh_pair_summary(fmdx_system_t *s)
{
    KDesc &D = s->D;
    int nk = D.nkinds;
    double rcmax2 = 0, rceam2 = 0;
    int cat = -2, has_eam = 0;
    for (int i = 0; i < nk; i++)
        for (int j = 0; j < nk; j++)
        {
            const PairParam &p = s->hp[i * nk + j];
            if (p.rc2 > rcmax2) rcmax2 = p.rc2;
            if (p.cat == FMDX_POT_EAM)
            {
                has_eam = 1;
                if (p.rc2 > rceam2) rceam2 = p.rc2;
            }
            if (cat == -2) cat = p.cat;
            else if (cat != p.cat) cat = -1;
        }
    D.rcmax2 = rcmax2;
    D.rceam2 = rceam2;
    D.has_eam = has_eam;
    D.single_cat = cat;
}

extern "C" int fmdx_set_pair(fmdx_system_t *s, int k1, int k2, int cat, double rc2, double p0, double p1, double p2)
{
    if (!s) return FMDX_EARG;
    int nk = s->D.nkinds;
    if (k1 < 0 || k2 < 0 || k1 >= nk || k2 >= nk) return FMDX_EARG;
    if (cat < FMDX_POT_NONE || cat > FMDX_POT_MORSE) return FMDX_EARG;
    PairParam p;
    memset(&p, 0, sizeof p);
    p.cat = cat;
    p.rc2 = rc2;
    p.p0 = p0; p.p1 = p1; p.p2 = p2;
    if (cat == FMDX_POT_EAM)
    {
        p.iloc = (int)p0;
        p.jloc = (int)p1;
    }
    s->hp[k1 * nk + k2] = p;
    if (cat == FMDX_POT_EAM) { int t = p.iloc; p.iloc = p.jloc; p.jloc = t; }
    s->hp[k2 * nk + k1] = p;
    refresh_pair_summary(s);
    CK(hipMemcpyAsync(s->dpairs, s->hp, sizeof(PairParam) * nk * nk, hipMemcpyHostToDevice, s->st));
    CK(hipStreamSynchronize(s->st));
    return ensure_n3(s, s->precision);
}

/* reference spline (y, yDD) on knots x_k = k*h -> per-interval coefficients in b = x/h - k */
static void spline_coeffs(const double *y, const double *ydd, int n, double h, double4 *out)
{
    const double h26 = h * h / 6.0;
    for (int k = 0; k < n - 1; k++)
    {
        double ylo = y[k], yhi = y[k + 1], dlo = ydd[k], dhi = ydd[k + 1];
        double4 q;
        q.x = ylo;
        q.y = (yhi - ylo) - h26 * (2.0 * dlo + dhi);
        q.z = 3.0 * h26 * dlo;
        q.w = h26 * (dhi - dlo);
        out[k] = q;
    }
}

/* Start (in r^2) of the steep run that ends a table at the cutoff: walking down from the
   interval holding rc^2, the intervals whose curvature |2 c2| + |6 c3| exceeds 10 x its median
   over [rc^2/4, rc^2].  A setfl table that is not zero at its cutoff ends in such a run (the two
   zero knots appended by eam.c:444-471 make a kink that the natural spline spreads over the
   intervals below it, decaying about 3.7x per r interval); a smooth table has none (+inf). */
static double steep_start(const double4 *c, long m, double dr2, double rc2)
{
    long kc = (long)(rc2 / dr2);
    if (kc > m - 1) kc = m - 1;
    long kq = (long)(0.25 * rc2 / dr2);
    if (kq > kc) kq = kc;
    std::vector<double> kap;
    kap.reserve(kc - kq + 1);
    for (long k = kq; k <= kc; k++) kap.push_back(fabs(2.0 * c[k].z) + fabs(6.0 * c[k].w));
    std::vector<double> srt(kap);
    std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
    const double lim = 10.0 * srt[srt.size() / 2];
    long k = kc;
    while (k >= kq && kap[k - kq] > lim) k--;
    return k == kc ? INFINITY : (double)(k + 1) * dr2;
}

extern "C" int fmdx_set_eam(fmdx_system_t *s, int nelem, int nrho, double drho, int nr2, double dr2, double rc2,
                            const double *F, const double *F_dd, const double *rho, const double *rho_dd,
                            const double *phi, const double *phi_dd, const int *elem_of_kind)
{
    if (!s || nelem < 1 || nelem > MAXE || nrho < 3 || nr2 < 3 || !(drho > 0) || !(dr2 > 0)) return FMDX_EARG;
    if (s->member) return FMDX_ESTATE;
    const long mr = nrho - 1, m2 = nr2 - 1;
    std::vector<double4> hF((size_t)nelem * mr), hR((size_t)nelem * m2), hP((size_t)nelem * nelem * m2);
    for (int e = 0; e < nelem; e++)
    {
        spline_coeffs(F + (size_t)e * nrho, F_dd + (size_t)e * nrho, nrho, drho, &hF[e * mr]);
        spline_coeffs(rho + (size_t)e * nr2, rho_dd + (size_t)e * nr2, nr2, dr2, &hR[e * m2]);
        for (int f = 0; f < nelem; f++)
            spline_coeffs(phi + ((size_t)e * nelem + f) * nr2, phi_dd + ((size_t)e * nelem + f) * nr2, nr2, dr2,
                          &hP[(e * nelem + f) * m2]);
    }
    std::vector<float4> fF(hF.size()), fR(hR.size()), fP(hP.size());
    auto cvt = [](const std::vector<double4> &a, std::vector<float4> &b) {
        for (size_t i = 0; i < a.size(); i++)
            b[i] = make_float4((float)a[i].x, (float)a[i].y, (float)a[i].z, (float)a[i].w);
    };
    cvt(hF, fF); cvt(hR, fR); cvt(hP, fP);
    /* element 0's force record per interval: the derivative coefficients of rho and phi
       first, the values last (a force pass without energy reads the first 48 / 24 bytes) */
    /* the derivative coefficients are stored as d/dx = (c1 + 2 c2 b + 3 c3 b^2) / h, that is
       (c1 / h, 2 c2 / h, 3 c3 / h), so that a derivative is two FMAs */
    std::vector<double4> hRF(2 * (size_t)m2);
    std::vector<float4> fRF(2 * (size_t)m2);
    const double ih = 1.0 / dr2;
    for (long k = 0; k < m2; k++)
    {
        const double4 r = hR[k], p = hP[k];
        const double r1 = r.y * ih, r2 = 2.0 * r.z * ih, r3 = 3.0 * r.w * ih;
        const double p1 = p.y * ih, p2 = 2.0 * p.z * ih, p3 = 3.0 * p.w * ih;
        hRF[2 * k] = make_double4(r1, r2, r3, p1);
        hRF[2 * k + 1] = make_double4(p2, p3, r.x, p.x);
        fRF[2 * k] = make_float4((float)r1, (float)r2, (float)r3, (float)p1);
        fRF[2 * k + 1] = make_float4((float)p2, (float)p3, (float)r.x, (float)p.x);
    }
    /* element 0's derivative knots of rho and phi on the reference's r^2 grid: D_k = h y'(x_k)
       (the slope in b = x/h - k, from interval k, or the end of the last one) and Y_k = y''_k h^2/6,
       so that within interval k  dy/db = D_k + b (6 Y_k + 3 b (Y_{k+1} - Y_k)).  The reference's
       own (y, y'') form would take the slope as y_{k+1} - y_k, which in FP32 cancels ~12 bits */
    std::vector<float4> fK((size_t)nr2);
    {
        const double h26 = dr2 * dr2 / 6.0;
        for (long k = 0; k < nr2; k++)
        {
            const long i = k < m2 ? k : m2 - 1;
            const double4 r = hR[i], p = hP[i];
            const double dr = k < m2 ? r.y : r.y + 2.0 * r.z + 3.0 * r.w;
            const double dp = k < m2 ? p.y : p.y + 2.0 * p.z + 3.0 * p.w;
            fK[k] = make_float4((float)dr, (float)dp, (float)(rho_dd[k] * h26), (float)(phi_dd[k] * h26));
        }
    }
    /* tables of an earlier call may still be read by work queued on the stream */
    CK(hipStreamSynchronize(s->st));
    const TabSizes tz{hF.size(), hR.size(), hP.size(), hRF.size(), fK.size()};
    carve_free(&s->tab_mem, [&](Carve &c) { tab_layout(s, c, tz); });
    if (carve_alloc(&s->tab_mem, [&](Carve &c) { tab_layout(s, c, tz); })) return FMDX_ENOMEM;
    CK(hipMemcpyAsync(s->Fc, hF.data(), hF.size() * sizeof(double4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->rhoc, hR.data(), hR.size() * sizeof(double4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->phic, hP.data(), hP.size() * sizeof(double4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->Fc32, fF.data(), fF.size() * sizeof(float4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->rhoc32, fR.data(), fR.size() * sizeof(float4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->phic32, fP.data(), fP.size() * sizeof(float4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->rf, hRF.data(), hRF.size() * sizeof(double4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->rf32, fRF.data(), fRF.size() * sizeof(float4), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(s->kn32, fK.data(), fK.size() * sizeof(float4), hipMemcpyHostToDevice, s->st));
    EamDev &E = s->D.eam;
    E.ne = nelem; E.nrho = nrho; E.nr2 = nr2;
    E.drho = drho; E.inv_drho = 1.0 / drho; E.dr2 = dr2; E.inv_dr2 = 1.0 / dr2; E.rc2 = rc2;
    /* FP32 kernels evaluate the pairs from rx2 up in FP64: the rounding band below rc^2 and the
       tables' steep runs (k_force.hip, exact pairs) */
    E.rx2 = rc2 * (1.0 - 1e-5);
    for (int e = 0; e < nelem; e++)
    {
        E.rx2 = fmin(E.rx2, steep_start(&hR[e * m2], m2, dr2, rc2));
        for (int f = 0; f < nelem; f++) E.rx2 = fmin(E.rx2, steep_start(&hP[(e * nelem + f) * m2], m2, dr2, rc2));
    }
    E.Fc = s->Fc; E.rhoc = s->rhoc; E.phic = s->phic;
    E.Fc32 = s->Fc32; E.rhoc32 = s->rhoc32; E.phic32 = s->phic32;
    E.rf = s->rf; E.rf32 = s->rf32; E.kn32 = s->kn32;
    /* the kernels run on the system stream: the tables have landed before any of them reads */
    CK(hipStreamSynchronize(s->st));
    /* hit words kept between the two EAM passes (one bit per lane and trip) */
    if (!s->D.hbits)
    {
        const size_t nw = (size_t)s->D.ncell * HBITS_WORDS * force_threads_per_cell();
        if (dalloc(&s->D.hbits, nw)) return FMDX_ENOMEM;
        /* never read before the density pass writes it; cleared all the same */
        CK(hipMemsetAsync(s->D.hbits, 0, nw * sizeof(unsigned long long), s->st));
        CK(hipStreamSynchronize(s->st));
    }
    for (int k = 0; k < s->D.nkinds; k++) s->D.elem[k] = elem_of_kind ? elem_of_kind[k] : -1;
    return FMDX_OK;
}

/* Slot capacities are multiples of 32, so that every cell's run of 4- and
   8-byte values starts on a 128-byte cache line: a copper cell's ~15 atoms
   then sit in one line of each array instead of straddling two. */
static unsigned line_cap(unsigned c)
{
    c = (c + 31) & ~31u;
    return c > 256 ? 256 : c;
}

static int choose_cap(unsigned maxc)
{
    return (int)line_cap(maxc + maxc / 2 + 8);
}

extern "C" int fmdx_upload(fmdx_system_t *s, unsigned natoms, const unsigned *cell_count, const double *x,
                           const double *v, const int *group, const unsigned *atomid, const unsigned *kind)
{
    if (!s || !cell_count) return FMDX_EARG;
    if (s->member) return FMDX_ESTATE;
    s->xq_fresh = 0;
    KDesc &D = s->D;
    std::vector<int> start(D.ncell + 1);
    unsigned maxc = 0;
    long acc = 0;
    for (int c = 0; c < D.ncell; c++)
    {
        start[c] = (int)acc;
        acc += cell_count[c];
        if (cell_count[c] > maxc) maxc = cell_count[c];
    }
    start[D.ncell] = (int)acc;
    if ((unsigned long)acc != natoms) return FMDX_EARG;
    if (maxc > 256) return FMDX_EARG;
    std::vector<uint8_t> k8(natoms ? natoms : 1);
    for (unsigned i = 0; i < natoms; i++)
    {
        unsigned k = kind ? kind[i] : 0;
        if ((int)k >= D.nkinds) return FMDX_EARG;
        k8[i] = (uint8_t)k;
    }
    int cap = choose_cap(maxc);
    hipStreamSynchronize(s->st);
    free_slots(D);
    int r = alloc_slots(D, cap, natoms + 64);
    if (r) return r;
    int *dstart;
    double *dx, *dv;
    int *dg;
    unsigned *da;
    uint8_t *dk;
    StageMem stage;
    auto stage_layout = [&](Carve &c) {
        c(&dstart, D.ncell + 1);
        c(&dx, 3 * (size_t)natoms); c(&dv, 3 * (size_t)natoms);
        c(&dg, natoms); c(&da, natoms); c(&dk, natoms);
    };
    if (carve_alloc(&stage.p, stage_layout)) return FMDX_ENOMEM;
    std::vector<int> g(natoms ? natoms : 1, 0);
    std::vector<unsigned> aid(natoms ? natoms : 1, 0);
    std::vector<double> vz(3 * (size_t)(natoms ? natoms : 1), 0.0);
    for (unsigned i = 0; i < natoms; i++)
    {
        if (group) g[i] = group[i];
        aid[i] = atomid ? atomid[i] : i;
    }
    CK(hipMemcpyAsync(dstart, start.data(), sizeof(int) * (D.ncell + 1), hipMemcpyHostToDevice, s->st));
    CK(hipMemcpyAsync(D.cnt, cell_count, sizeof(int) * D.ncell, hipMemcpyHostToDevice, s->st));
    if (natoms)
    {
        CK(hipMemcpyAsync(dx, x, sizeof(double) * 3 * natoms, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(dv, v ? v : vz.data(), sizeof(double) * 3 * natoms, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(dg, g.data(), sizeof(int) * natoms, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(da, aid.data(), sizeof(unsigned) * natoms, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(dk, k8.data(), natoms, hipMemcpyHostToDevice, s->st));
    }
    /* binning moves atoms into slots that held none; their force words must
       read zero until the next force evaluation, as for a fresh host cell */
    const long nslot = (long)D.ncell * D.cap;
    CK(launch_fill(D.fx, nslot, 0.0, s->st));
    CK(launch_fill(D.fy, nslot, 0.0, s->st));
    CK(launch_fill(D.fz, nslot, 0.0, s->st));
    CK(launch_fill(D.aux, nslot, 0.0, s->st));
    CK(launch_scatter(D, s->st, dstart, dx, dv, dg, da, dk));
    unsigned ctr[NCOUNTERS] = {0, 0, maxc};
    CK(hipMemcpyAsync(D.counters, ctr, sizeof ctr, hipMemcpyHostToDevice, s->st));
    CK(hipStreamSynchronize(s->st));
    s->natoms_uploaded = natoms;
    s->uploaded = 1;
    return ensure_n3(s, s->precision);
}

extern "C" int fmdx_natoms(fmdx_system_t *s, unsigned *n)
{
    if (!s || !n) return FMDX_EARG;
    unsigned ctr[8];
    CK(hipMemcpyAsync(ctr, s->D.counters, sizeof ctr, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    *n = s->natoms_uploaded - ctr[1];
    return FMDX_OK;
}

extern "C" int fmdx_ncells(fmdx_system_t *s, unsigned *n)
{
    if (!s || !n) return FMDX_EARG;
    *n = (unsigned)s->D.ncell;
    return FMDX_OK;
}

extern "C" int fmdx_download(fmdx_system_t *s, unsigned *natoms, unsigned *cell_count, double *x, double *v,
                             double *f, double *aux, int *group, unsigned *atomid, unsigned *kind)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    KDesc &D = s->D;
    std::vector<unsigned> cnt(D.ncell);
    CK(hipMemcpyAsync(cnt.data(), D.cnt, sizeof(int) * D.ncell, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    std::vector<int> start(D.ncell + 1);
    long acc = 0;
    for (int c = 0; c < D.ncell; c++) { start[c] = (int)acc; acc += cnt[c]; }
    start[D.ncell] = (int)acc;
    unsigned n = (unsigned)acc;
    if (natoms) *natoms = n;
    if (cell_count) memcpy(cell_count, cnt.data(), sizeof(unsigned) * D.ncell);
    if (!x && !v && !f && !aux && !group && !atomid && !kind) return FMDX_OK;
    int *dstart;
    double *dx, *dv, *df, *daux;
    int *dg;
    unsigned *da;
    uint8_t *dk;
    StageMem stage;
    auto stage_layout = [&](Carve &c) {
        c(&dstart, D.ncell + 1);
        c(&dx, 3 * (size_t)n); c(&dv, 3 * (size_t)n); c(&df, 3 * (size_t)n);
        c(&daux, n); c(&dg, n); c(&da, n); c(&dk, n);
    };
    if (carve_alloc(&stage.p, stage_layout)) return FMDX_ENOMEM;
    CK(hipMemcpyAsync(dstart, start.data(), sizeof(int) * (D.ncell + 1), hipMemcpyHostToDevice, s->st));
    CK(launch_gather(D, s->st, dstart, dx, dv, df, daux, dg, da, dk));
    if (n)
    {
        if (x) CK(hipMemcpyAsync(x, dx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->st));
        if (v) CK(hipMemcpyAsync(v, dv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->st));
        if (f) CK(hipMemcpyAsync(f, df, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->st));
        if (aux) CK(hipMemcpyAsync(aux, daux, sizeof(double) * n, hipMemcpyDeviceToHost, s->st));
        if (group) CK(hipMemcpyAsync(group, dg, sizeof(int) * n, hipMemcpyDeviceToHost, s->st));
        if (atomid) CK(hipMemcpyAsync(atomid, da, sizeof(unsigned) * n, hipMemcpyDeviceToHost, s->st));
    }
    std::vector<uint8_t> k8(n ? n : 1);
    if (kind && n) CK(hipMemcpyAsync(k8.data(), dk, n, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    if (kind)
        for (unsigned i = 0; i < n; i++) kind[i] = k8[i];
    return FMDX_OK;
}

TtmDev ttm_dev(fmdx_system_t *s, int on)
{
    TtmDev T;
    memset(&T, 0, sizeof T);
    if (on && s->ttm_on)
    {
        T.on = 1;
        T.ncz = s->T.ncz;
        T.tcellh_z = s->T.tcellh_z;
        T.xi = s->T.xi;
        T.vcm = s->T.vcm;
        T.ncx = s->T.ncx;
        T.ncy = s->T.ncy;
        T.tcellh_x = s->T.tcellh_x;
        T.tcellh_y = s->T.tcellh_y;
    }
    return T;
}

extern "C" int fmdx_update_forces(fmdx_system_t *s, int active_group, int ttm_friction)
{
    return fmdx_update_forces_ex(s, active_group, ttm_friction, FMDX_FORCES_ENERGY);
}

/* Allocated as soon as a system with slots qualifies (upload, precision, potentials,
   capacity), so that the allocation-free entry point finds them and takes the same
   kernels as fmdx_update_forces; the regular entry point and the batch engine also
   allocate on first need. */
int ensure_n3(fmdx_system_t *s, int precision)
{
    if (!s->uploaded || !n3_wanted(s->D, precision) || s->D.np) return FMDX_OK;
    return dalloc(&s->D.np, (size_t)39 * s->D.ncell * s->D.cap);
}

extern "C" int fmdx_update_forces_ex(fmdx_system_t *s, int active_group, int ttm_friction, int flags)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    if (s->D.has_eam && !s->rhoc) return FMDX_ESTATE;
    if (flags & ~FMDX_FORCES_ENERGY) return FMDX_EARG;
    int r = ensure_n3(s, s->precision);
    if (r) return r;
    s->D.ag = active_group;
    s->D.energy = (flags & FMDX_FORCES_ENERGY) ? 1 : 0;
    s->D.ttm = ttm_dev(s, ttm_friction);
    /* the FP32 cell-relative positions are formed here unless the integrator step just
       before left them in place */
    const bool relpos = !s->xq_fresh;
    s->xq_fresh = 0;
    if (s->profile && s->prof_n < NPROF)
    {
        hipEvent_t *e = &s->pev[3 * s->prof_n];
        CK(hipEventRecord(e[0], s->st));
        CK(launch_pass1(s->D, s->st, s->precision, relpos));
        CK(hipEventRecord(e[1], s->st));
        CK(launch_pass0(s->D, s->st, s->precision));
        CK(hipEventRecord(e[2], s->st));
        if (s->D.energy) CK(launch_reduce_pe(s->D, s->st));
        s->prof_n++;
    }
    else
        CK(launch_forces(s->D, s->st, s->precision, relpos));
    return FMDX_OK;
}

extern "C" int fmdx_vv_start(fmdx_system_t *s, int active_group, double dt, int thermostat, double tau, double T0)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    set_step(s->D, s->mass, active_group, dt, thermostat, tau, T0);
    s->xq_fresh = 0;
    CK(launch_vv_start(s->D, s->st));
    s->xq_fresh = s->D.xq_keep;
    return FMDX_OK;
}

extern "C" int fmdx_vv_finish(fmdx_system_t *s, int active_group, double dt)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    set_step(s->D, s->mass, active_group, dt, 0, 1.0, 0.0);
    CK(launch_vv_finish(s->D, s->st));
    return FMDX_OK;
}

extern "C" int fmdx_se_step(fmdx_system_t *s, int active_group, double dt)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    set_step(s->D, s->mass, active_group, dt, 0, 1.0, 0.0);
    s->xq_fresh = 0;
    CK(launch_se_step(s->D, s->st));
    s->xq_fresh = s->D.xq_keep;
    return FMDX_OK;
}

extern "C" int fmdx_refresh(fmdx_system_t *s)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    s->xq_fresh = 0;
    CK(launch_flag_only(s->D, s->st));
    return FMDX_OK;
}

extern "C" int fmdx_thermo(fmdx_system_t *s, int active_group)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    s->D.ag = active_group;
    CK(launch_thermo(s->D, s->st));
    return FMDX_OK;
}

extern "C" int fmdx_add_velocity(fmdx_system_t *s, int group, double vx, double vy, double vz)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    CK(launch_add_velocity(s->D, s->st, group, vx, vy, vz));
    return FMDX_OK;
}

extern "C" int fmdx_read_thermo(fmdx_system_t *s, fmdx_thermo_t *out)
{
    if (!s || !out) return FMDX_EARG;
    double th[16];
    unsigned ctr[NCOUNTERS];
    CK(hipMemcpyAsync(th, s->D.thermo, sizeof th, hipMemcpyDeviceToHost, s->st));
    CK(hipMemcpyAsync(ctr, s->D.counters, sizeof ctr, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    out->pe = th[TH_PE];
    out->embed = th[TH_EMB];
    out->mv2 = th[TH_MV2];
    out->ke = th[TH_KE];
    out->temperature = th[TH_T];
    out->momentum[0] = th[TH_PX];
    out->momentum[1] = th[TH_PY];
    out->momentum[2] = th[TH_PZ];
    out->n_active = th[TH_N];
    out->natoms = s->natoms_uploaded - ctr[1];
    out->max_cell_count = ctr[2];
    out->err = ctr[3];
    out->cap = (unsigned)s->D.cap;
    out->bad_force_cell = ctr[4];
    out->bad_position_cell = ctr[5];
    out->replay_mismatch = ctr[6];
    out->exact_unmatched = ctr[7];
    out->audit_atoms = ctr[8];
    out->audit_max_cell = ctr[9];
    out->audit_nonfinite = ctr[10];
    return FMDX_OK;
}

extern "C" int fmdx_audit(fmdx_system_t *s)
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    CK(launch_audit(s->D, s->st));
    return FMDX_OK;
}

extern "C" int fmdx_clear_errors(fmdx_system_t *s)
{
    if (!s) return FMDX_EARG;
    unsigned z = 0;
    CK(hipMemcpyAsync(&s->D.counters[3], &z, sizeof z, hipMemcpyHostToDevice, s->st));
    CK(hipStreamSynchronize(s->st));
    return FMDX_OK;
}

/* device bytes held by a system (slot layout, per-cell arrays, tables) */
extern "C" int fmdx_memory(fmdx_system_t *s, size_t *bytes)
{
    if (!s || !bytes) return FMDX_EARG;
    const KDesc &D = s->D;
    const size_t ns = (size_t)D.ncell * D.cap;
    const size_t per_slot = 10 * sizeof(double) + 3 * sizeof(float) + 2 * sizeof(int) + sizeof(unsigned) + 2 +
                            sizeof(unsigned long long) + sizeof(int);
    const size_t nmov = s->natoms_uploaded + 64;
    size_t b = s->uploaded ? ns * per_slot + nmov * (6 * sizeof(double) + 9) : 0;
    b += (size_t)D.ncell * (6 * sizeof(int) + 2 * sizeof(double) + sizeof(unsigned long long));
    if (D.hbits) b += (size_t)D.ncell * HBITS_WORDS * force_threads_per_cell() * sizeof(unsigned long long);
    if (s->rhoc) b += (size_t)(D.eam.nrho + 6 * D.eam.nr2) * (sizeof(double4) + sizeof(float4));
    if (D.np) b += ns * 39 * sizeof(float);
    *bytes = b;
    return FMDX_OK;
}

extern "C" int fmdx_sync(fmdx_system_t *s)
{
    if (!s) return FMDX_EARG;
    CK(hipStreamSynchronize(s->st));
    return FMDX_OK;
}

extern "C" int fmdx_set_capacity(fmdx_system_t *s, unsigned cap)
{
    if (!s || !s->uploaded || cap < 8 || cap > 256) return FMDX_EARG;
    if (s->member) return FMDX_ESTATE;
    unsigned ctr[8];
    CK(hipMemcpyAsync(ctr, s->D.counters, sizeof ctr, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    if (cap < ctr[2]) return FMDX_EARG;
    KDesc N = s->D;
    N.np = nullptr;   /* the partner sums of the new layout are allocated below */
    int r = alloc_slots(N, (int)cap, s->natoms_uploaded + 64);
    if (r) return r;
    const long nslot = (long)N.ncell * N.cap;
    CK(launch_fill(N.fx, nslot, 0.0, s->st));
    CK(launch_fill(N.fy, nslot, 0.0, s->st));
    CK(launch_fill(N.fz, nslot, 0.0, s->st));
    CK(launch_fill(N.aux, nslot, 0.0, s->st));
    CK(launch_relayout(s->D, N, s->st));
    CK(hipStreamSynchronize(s->st));
    free_slots(s->D);
    s->D = N;
    s->xq_fresh = 0;   /* the new layout's FP32 positions are formed at the next force call */
    return ensure_n3(s, s->precision);
}

/* ------------------------------------------------------------------ TTM */

template <class T>
static int ttm_alloc(fmdx_system_t *s, T **p, size_t n)
{
    int r = dalloc(p, n);
    if (r) return r;
    s->ttm_mem[s->nttm_mem++] = (void *)*p;
    hipMemsetAsync(*p, 0, n * sizeof(T), s->st);
    return FMDX_OK;
}

extern "C" int fmdx_ttm_setup(fmdx_system_t *s, const fmdx_ttm_t *c, const double *Te_init)
{
    if (!s || !c || c->ncz < 1 || c->ratio < 1) return FMDX_EARG;
    if (c->type != 1 && c->type != 2) return FMDX_EARG;
    const int ncx = c->ncx > 0 ? c->ncx : 1, ncy = c->ncy > 0 ? c->ncy : 1;
    const int dim3 = ncx * ncy > 1 || c->kmode != 0 || c->solver3d;
    if (!dim3 && c->ncz > FMDX_TTM_MAX_CELLS) return FMDX_EARG;
    if (dim3 && (c->type != 1 || (long)ncx * ncy * (c->ncz + 2) > FMDX_TTM_MAX_TCELLS)) return FMDX_EARG;
    if (c->type == 2 && (c->nC < 2 || c->nG < 2 || !c->T_C || !c->Ct || !c->C_DD || !c->T_G || !c->Gt || !c->G_DD))
        return FMDX_EARG;
    for (int i = 0; i < s->nttm_mem; i++) hipFree(s->ttm_mem[i]);
    s->nttm_mem = 0;
    TtmState &S = s->T;
    memset(&S, 0, sizeof S);
    S.type = c->type; S.ncz = c->ncz; S.ratio = c->ratio;
    S.ncx = ncx; S.ncy = ncy; S.nzg = c->ncz + 2; S.n = ncx * ncy * S.nzg;
    S.dim3 = dim3; S.kmode = c->kmode;
    S.tcellh_x = c->tcellh_x; S.tcellh_y = c->tcellh_y;
    S.tcellh_z = c->tcellh_z; S.tvol = c->tcell_volume;
    S.C_gamma = c->C_gamma; S.K = c->K; S.G = c->G;
    S.Kzh_A = c->Kzh_A; S.Kzh_B = c->Kzh_B; S.Kzh_v2 = c->Kzh_v2;
    S.nC = c->nC; S.nG = c->nG;
    S.laser_factor = c->laser_factor; S.laser_absdepth_inv = c->laser_absdepth_inv;
    S.laser_2sig2_inv = c->laser_2sig2_inv; S.laser_t0 = c->laser_t0; S.laser_tdiff = c->laser_tdiff;
    S.frontsurf = c->frontsurf; S.min_atoms = c->min_atoms;
    int r = 0;
    r |= ttm_alloc(s, &S.mass, S.n); r |= ttm_alloc(s, &S.vcm, 3 * S.n); r |= ttm_alloc(s, &S.Ti, S.n);
    r |= ttm_alloc(s, &S.Te, S.n); r |= ttm_alloc(s, &S.Te2, S.n); r |= ttm_alloc(s, &S.xi, S.n);
    r |= ttm_alloc(s, &S.Ke, S.n); r |= ttm_alloc(s, &S.Ce, S.n); r |= ttm_alloc(s, &S.Gc, S.n);
    r |= ttm_alloc(s, &S.num, S.n);
    if (r) return r;
    if (c->type == 2)
    {
        double *tc, *ct, *cdd, *tg, *gt, *gdd;
        r |= ttm_alloc(s, &tc, c->nC); r |= ttm_alloc(s, &ct, c->nC); r |= ttm_alloc(s, &cdd, c->nC);
        r |= ttm_alloc(s, &tg, c->nG); r |= ttm_alloc(s, &gt, c->nG); r |= ttm_alloc(s, &gdd, c->nG);
        if (r) return r;
        CK(hipMemcpyAsync(tc, c->T_C, sizeof(double) * c->nC, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(ct, c->Ct, sizeof(double) * c->nC, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(cdd, c->C_DD, sizeof(double) * c->nC, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(tg, c->T_G, sizeof(double) * c->nG, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(gt, c->Gt, sizeof(double) * c->nG, hipMemcpyHostToDevice, s->st));
        CK(hipMemcpyAsync(gdd, c->G_DD, sizeof(double) * c->nG, hipMemcpyHostToDevice, s->st));
        S.T_C = tc; S.Ct = ct; S.C_DD = cdd; S.T_G = tg; S.Gt = gt; S.G_DD = gdd;
    }
    if (Te_init) CK(hipMemcpyAsync(S.Te, Te_init, sizeof(double) * S.n, hipMemcpyHostToDevice, s->st));
    CK(hipStreamSynchronize(s->st));
    s->ttm_on = 1;
    return FMDX_OK;
}

extern "C" int fmdx_ttm_set_frontsurf(fmdx_system_t *s, double frontsurf, unsigned min_atoms)
{
    if (!s || !s->ttm_on) return FMDX_ESTATE;
    s->T.frontsurf = frontsurf;
    s->T.min_atoms = min_atoms;
    return FMDX_OK;
}

extern "C" int fmdx_ttm_fields(fmdx_system_t *s)
{
    if (!s || !s->ttm_on || !s->uploaded) return FMDX_ESTATE;
    CK(launch_ttm_fields(s->D, s->T, s->st));
    return FMDX_OK;
}

extern "C" int fmdx_ttm_update(fmdx_system_t *s, double time, double dt)
{
    if (!s || !s->ttm_on || !s->uploaded) return FMDX_ESTATE;
    CK(launch_ttm_fields(s->D, s->T, s->st));
    CK(launch_ttm_presolve(s->T, &s->D.counters[3], s->st));
    CK(launch_ttm_solve(s->T, time, dt, &s->D.counters[3], s->st));
    return FMDX_OK;
}

extern "C" int fmdx_ttm_presolve(fmdx_system_t *s)
{
    if (!s || !s->ttm_on) return FMDX_ESTATE;
    CK(launch_ttm_presolve(s->T, &s->D.counters[3], s->st));
    return FMDX_OK;
}

extern "C" int fmdx_ttm_solve(fmdx_system_t *s, double time, double dt)
{
    if (!s || !s->ttm_on) return FMDX_ESTATE;
    CK(launch_ttm_solve(s->T, time, dt, &s->D.counters[3], s->st));
    return FMDX_OK;
}

extern "C" int fmdx_ttm_download(fmdx_system_t *s, double *Te, double *Ti, double *xi, double *vcm, unsigned *num,
                                 double *mass)
{
    if (!s || !s->ttm_on) return FMDX_ESTATE;
    const TtmState &S = s->T;
    if (Te) CK(hipMemcpyAsync(Te, S.Te, sizeof(double) * S.n, hipMemcpyDeviceToHost, s->st));
    if (Ti) CK(hipMemcpyAsync(Ti, S.Ti, sizeof(double) * S.n, hipMemcpyDeviceToHost, s->st));
    if (xi) CK(hipMemcpyAsync(xi, S.xi, sizeof(double) * S.n, hipMemcpyDeviceToHost, s->st));
    if (vcm) CK(hipMemcpyAsync(vcm, S.vcm, sizeof(double) * 3 * S.n, hipMemcpyDeviceToHost, s->st));
    if (num) CK(hipMemcpyAsync(num, S.num, sizeof(unsigned) * S.n, hipMemcpyDeviceToHost, s->st));
    if (mass) CK(hipMemcpyAsync(mass, S.mass, sizeof(double) * S.n, hipMemcpyDeviceToHost, s->st));
    CK(hipStreamSynchronize(s->st));
    return FMDX_OK;
}

extern "C" int fmdx_ttm_upload_te(fmdx_system_t *s, const double *Te)
{
    if (!s || !s->ttm_on || !Te) return FMDX_ESTATE;
    CK(hipMemcpyAsync(s->T.Te, Te, sizeof(double) * s->T.n, hipMemcpyHostToDevice, s->st));
    CK(hipStreamSynchronize(s->st));
    return FMDX_OK;
}

extern "C" int fmdx_find_limits(fmdx_system_t *s, int group, double lo[3], double hi[3])
{
    if (!s || !s->uploaded) return FMDX_ESTATE;
    unsigned n = 0;
    int r = fmdx_download(s, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (r) return r;
    std::vector<double> x(3 * (size_t)(n ? n : 1));
    std::vector<int> g(n ? n : 1);
    r = fmdx_download(s, &n, nullptr, x.data(), nullptr, nullptr, nullptr, g.data(), nullptr, nullptr);
    if (r) return r;
    for (int d = 0; d < 3; d++) { lo[d] = 1.7976931348623157e308; hi[d] = 2.2250738585072014e-308; }
    for (unsigned i = 0; i < n; i++)
        if (group == -1 || g[i] == group)
            for (int d = 0; d < 3; d++)
            {
                if (x[3 * i + d] < lo[d]) lo[d] = x[3 * i + d];
                if (x[3 * i + d] > hi[d]) hi[d] = x[3 * i + d];
            }
    return FMDX_OK;
}

extern "C" int fmdx_event_record(fmdx_system_t *s, int slot)
{
    if (!s || slot < 0 || slot >= 14) return FMDX_EARG;
    CK(hipEventRecord(s->ev[slot], s->st));
    return FMDX_OK;
}

extern "C" int fmdx_event_elapsed_ms(fmdx_system_t *s, int a, int b, float *ms)
{
    if (!s || a < 0 || b < 0 || a >= 14 || b >= 14) return FMDX_EARG;
    CK(hipEventSynchronize(s->ev[b]));
    CK(hipEventElapsedTime(ms, s->ev[a], s->ev[b]));
    return FMDX_OK;
}

extern "C" int fmdx_profile_enable(fmdx_system_t *s, int on)
{
    if (!s) return FMDX_EARG;
    s->profile = on;
    s->prof_n = 0;
    return FMDX_OK;
}

/* ms[0] = pass-1 (EAM density) kernel time, ms[1] = pass-0 (force) kernel time,
   ms[2] = both, summed over the profiled force calls; count = number of calls */
extern "C" int fmdx_profile_read(fmdx_system_t *s, double *ms, unsigned *count)
{
    if (!s || !ms) return FMDX_EARG;
    for (int i = 0; i < 8; i++) ms[i] = 0.0;
    CK(hipStreamSynchronize(s->st));
    for (unsigned k = 0; k < s->prof_n; k++)
    {
        float a = 0, b = 0;
        CK(hipEventElapsedTime(&a, s->pev[3 * k], s->pev[3 * k + 1]));
        CK(hipEventElapsedTime(&b, s->pev[3 * k + 1], s->pev[3 * k + 2]));
        ms[0] += a;
        ms[1] += b;
        ms[2] += a + b;
    }
    if (count) *count = s->prof_n;
    return FMDX_OK;
}
