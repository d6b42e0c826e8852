/*
 * k_force.hip -- 27-neighbour-cell pair kernels for LJ, Morse and two-pass EAM.
 *
 * One 64-lane wavefront (one workgroup) owns one home cell.
 *  1. Staging.  The atoms of the 27 neighbour cells are copied into an LDS
 *     candidate tile with coalesced SoA loads.  Periodic images are shifted
 *     into the home cell's frame on the way in, as the reference's ghost
 *     copies are (md-ghost.c:85-107).  Candidates outside the active group
 *     are not staged, and the padding of the last trip lies where it never
 *     matches.  A list longer than one tile is processed tile by tile.
 *  2. Scan.  Lanes are split into (home atom a, stripe s): with n home atoms
 *     each atom gets S = 64/n lanes and lane (a,s) tests candidates s, s+S,
 *     s+2S, ... of the tile.  The trip count is wave-uniform, and every 64
 *     trips leave one 64-bit word of hits per lane.  The spline / exp work
 *     then runs over the set bits of that word, so it is never executed on
 *     behalf of a lane without a hit.
 *  3. Replay.  The EAM density pass stores its hit words; the force pass of
 *     the same step stages the same tiles and walks those words instead of
 *     scanning again.  Cells whose words do not fit are rescanned.
 * The partial sums of the S stripes are combined in a fixed order, so every
 * atom's sum is deterministic.
 *
 * Full shell: each atom accumulates only its own force / density, which
 * removes the reference's Newton-3 writes into neighbour cells (the OpenMP
 * race of lj.c:37-87, eam.c:38-142) and the ghost fold-back
 * (md-ghost.c:788-825).
 *
 * Pair set and formulas:
 *   LJ      lj.h:26-58        F += 12*(4eps)/r^2 * ((s/r)^12 - (s/r)^6/2) r,   U = 4eps((s/r)^12-(s/r)^6)
 *   Morse   morse.h:26-58     F += 2 a D0/r (e^2 - e) r, e = exp(-a(r-r0)),    U = D0(e^2 - 2e)
 *   EAM p1  eam.h:270-379     rho_i = sum_j rho_elem(j)(r^2)   then F'(rho), F(rho)  eam.c:160-223
 *   EAM p0  eam.h:29-228      F_i -= 2(F'_i rho'_elem(j) + F'_j rho'_elem(i) + phi'_ij) r
 * with every tabulated function a natural cubic spline on the reference's
 * uniform r^2 (or rho) grid, stored as per-interval polynomial coefficients.
 * The pair energy of a pair is accumulated half on each side.
 *
 * Two precisions share this code: R = double (positions, arithmetic and sums
 * in FP64) and R = float (FP32 positions relative to the home cell origin,
 * FP32 pair arithmetic, Kahan-compensated FP32 sums, FP64 results).
 */
#include "fmdx_internal.h"

#define WG 64                             /* lanes per wavefront */
#define NT (WG * FMDX_WAVES_PER_CELL)     /* threads per home cell (one workgroup) */
#define NOKIND 0xFF    /* padding or a candidate outside the active group */

/* pair-category specialisations chosen on the host */
enum
{
    CAT_GEN = 0,     /* per-pair dispatch through the pair table */
    CAT_EAM1 = 1,    /* EAM everywhere, one element: interleaved rho/phi records */
    CAT_LJ = 2,      /* one kind, LJ */
    CAT_MORSE = 3    /* one kind, Morse */
};

/* Candidates per LDS tile.  The tile sets a workgroup's LDS and with it the
   workgroups per CU; a cell's candidates beyond one tile take further rounds.
   A round stops staging once fewer than 64 places are left, and about 49 of
   every 64 staged candidates survive the culling, so 224 (FP64) still takes
   the ~300 survivors of a copper cell in two rounds of ~196, as 256 does,
   with less LDS: the FP64 force pass needs 9.7 KB, which lets 16 workgroups
   share a CU, as many as its registers allow.  Measured on the 4M-atom EAM
   step (FP64 / mixed): 224/384 4.27/3.33 ms, 240/416 4.34/3.34, 256/512
   4.39/3.41.  The values in use come from fmdx_tuned.h. */
#include "fmdx_tuned.h"
template <typename R> struct Prec;
template <> struct Prec<double>
{
    enum { CT = FMDX_TILE64 };
    typedef double4 Q;
    static __device__ __forceinline__ const double4 *rho(const EamDev &E) { return E.rhoc; }
    static __device__ __forceinline__ const double4 *phi(const EamDev &E) { return E.phic; }
    static __device__ __forceinline__ const void *rf(const EamDev &E) { return E.rf; }
    /* the slot's position and (FP) F' */
    template <bool FP>
    static __device__ __forceinline__ void load(const KDesc &D, long s, double &x, double &y, double &z, double &f)
    {
        x = D.x[s]; y = D.y[s]; z = D.z[s];
        if (FP) f = D.aux[s];
    }
};
template <> struct Prec<float>
{
    enum { CT = FMDX_TILE32 };   /* the generic kernel; the others use FMDX_TILE32_EAM / _LJ / _MORSE */
    typedef float4 Q;
    static __device__ __forceinline__ const float4 *rho(const EamDev &E) { return E.rhoc32; }
    static __device__ __forceinline__ const float4 *phi(const EamDev &E) { return E.phic32; }
    static __device__ __forceinline__ const void *rf(const EamDev &E) { return E.rf32; }
    /* the slot's packed record: one 16-byte load for the position and F' */
    template <bool FP>
    static __device__ __forceinline__ void load(const KDesc &D, long s, float &x, float &y, float &z, float &f)
    {
        const float4 q = D.xq[s];
        x = q.x; y = q.y; z = q.z;
        if (FP) f = q.w;
    }
};

/* LDS candidate tile (positions in the home frame, F' of the candidate, and
   the kind where the generic kernel dispatches on it: KD); N covers the tile
   plus the padding of the last scan trip */
template <typename R, bool FP, bool KD, int CT> struct Tile;
struct Rec3 { double x, y, z; };
template <bool FP> struct RecOf { typedef double4 T; };
template <> struct RecOf<false> { typedef Rec3 T; };
template <bool FP, bool KD, int CT> struct Tile<double, FP, KD, CT>
{
    enum { N = CT + NT };
    typename RecOf<FP>::T p[N];   /* x, y, z (, F') */
    uint8_t k[KD ? N : 1];
    __device__ __forceinline__ void put(int m, double x, double y, double zz, double f, int kk)
    {
        p[m].x = x; p[m].y = y; p[m].z = zz;
        if constexpr (FP) p[m].w = f;
        if constexpr (KD) k[m] = (uint8_t)kk;
    }
    __device__ __forceinline__ void pos(int m, double &x, double &y, double &zz) const
    {
        x = p[m].x; y = p[m].y; zz = p[m].z;
    }
    __device__ __forceinline__ const typename RecOf<FP>::T *at(int m) const { return p + m; }
    __device__ __forceinline__ void load(int m, double &x, double &y, double &zz, double &f) const
    {
        x = p[m].x; y = p[m].y; zz = p[m].z;
        if constexpr (FP) f = p[m].w;
        else f = 0.0;
    }
    /* entry t of a lane's word: the lane's first entry m0, stripes S entries apart, as a
       byte offset t * S * size (one 24-bit multiply-add per entry) */
    __device__ __forceinline__ unsigned stride(int S) const { return (unsigned)S * (unsigned)sizeof(p[0]); }
    __device__ __forceinline__ void load(int m0, unsigned st, int t, double &x, double &y, double &zz, double &f) const
    {
        const auto &e = *reinterpret_cast<const typename RecOf<FP>::T *>(
            reinterpret_cast<const char *>(p + m0) + __umul24((unsigned)t, st));
        x = e.x; y = e.y; zz = e.z;
        if constexpr (FP) f = e.w;
        else f = 0.0;
    }
};
template <bool FP, bool KD, int CT> struct Tile<float, FP, KD, CT>
{
    enum { N = CT + NT };
    float4 p[N];
    unsigned short g[N];   /* the candidate's place in the 27-cell list, for the exact FP64 pairs */
    uint8_t k[KD ? N : 1];
    __device__ __forceinline__ void put(int m, float x, float y, float zz, float f, int kk)
    {
        p[m] = make_float4(x, y, zz, f);
        if constexpr (KD) k[m] = (uint8_t)kk;
    }
    __device__ __forceinline__ void pos(int m, float &x, float &y, float &zz) const
    {
        const float4 a = p[m];
        x = a.x; y = a.y; zz = a.z;
    }
    __device__ __forceinline__ void load(int m, float &x, float &y, float &zz, float &f) const
    {
        const float4 a = p[m];
        x = a.x; y = a.y; zz = a.z; f = a.w;
    }
    __device__ __forceinline__ const float4 *at(int m) const { return p + m; }
    __device__ __forceinline__ unsigned stride(int S) const { return (unsigned)S * (unsigned)sizeof(float4); }
    __device__ __forceinline__ void load(int m0, unsigned st, int t, float &x, float &y, float &zz, float &f) const
    {
        const float4 a = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(p + m0) + __umul24((unsigned)t, st));
        x = a.x; y = a.y; zz = a.z; f = a.w;
    }
};

/* Image shift of each of the 27 neighbour cells, added to the staged
   positions.  FP64 keeps the wrap (-1, 0 or +1 per dimension) and multiplies
   it by the box length when staging, which gives the same x_j - L, x_j or
   x_j + L as a stored shift; FP32 keeps the origin of the neighbour's image
   relative to the home cell origin, formed in FP64 and rounded once. */
template <typename R> struct Shifts;
template <> struct Shifts<double>
{
    signed char w[3][27];
    __device__ __forceinline__ void set(const KDesc &D, int n, const int (&wr)[3], const int (&dj)[3], const int (&hc)[3])
    {
        (void)D; (void)dj; (void)hc;
        w[0][n] = (signed char)wr[0]; w[1][n] = (signed char)wr[1]; w[2][n] = (signed char)wr[2];
    }
    __device__ __forceinline__ double sx(const KDesc &D, int n) const { return (double)w[0][n] * D.L0; }
    __device__ __forceinline__ double sy(const KDesc &D, int n) const { return (double)w[1][n] * D.L1; }
    __device__ __forceinline__ double sz(const KDesc &D, int n) const { return (double)w[2][n] * D.L2; }
};
template <> struct Shifts<float>
{
    float s[3][27];
    signed char w[3][27];   /* the wrap, for the exact FP64 pairs */
    __device__ __forceinline__ void set(const KDesc &D, int n, const int (&wr)[3], const int (&dj)[3], const int (&hc)[3])
    {
        s[0][n] = (float)((dj[0] * D.h0 + wr[0] * D.L0) - hc[0] * D.h0);
        s[1][n] = (float)((dj[1] * D.h1 + wr[1] * D.L1) - hc[1] * D.h1);
        s[2][n] = (float)((dj[2] * D.h2 + wr[2] * D.L2) - hc[2] * D.h2);
        w[0][n] = (signed char)wr[0]; w[1][n] = (signed char)wr[1]; w[2][n] = (signed char)wr[2];
    }
    __device__ __forceinline__ float sx(const KDesc &D, int n) const { (void)D; return s[0][n]; }
    __device__ __forceinline__ float sy(const KDesc &D, int n) const { (void)D; return s[1][n]; }
    __device__ __forceinline__ float sz(const KDesc &D, int n) const { (void)D; return s[2][n]; }
};

/* the 32-bit word holding the sign bit of a value, in bit 31 */
static __device__ __forceinline__ unsigned sign_word(float v) { return __builtin_bit_cast(unsigned, v); }
static __device__ __forceinline__ unsigned sign_word(double v)
{
    return (unsigned)(__builtin_bit_cast(unsigned long long, v) >> 32);
}

/* accumulators: plain FP64, or Kahan-compensated FP32 */
struct AccD
{
    double v;
    __device__ __forceinline__ void zero() { v = 0.0; }
    __device__ __forceinline__ void add(double x) { v += x; }
    __device__ __forceinline__ double get() const { return v; }
};
/* FP32 compensated (Kahan) sum.  The EAM flush loops add the two to four pair
   terms of one iteration in plain FP32 first and compensate once per
   iteration, the LJ and Morse flushes once per hit word (up to 64 trips):
   one rounding per term of a short partial sum (relative 6e-8 of it)
   instead of three extra operations per term. */
struct AccF
{
    float s, c;
    __device__ __forceinline__ void zero() { s = 0.0f; c = 0.0f; }
    __device__ __forceinline__ void add(float v)
    {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double get() const { return (double)s - (double)c; }
};
template <typename R> struct AccOf;
template <> struct AccOf<double> { typedef AccD T; };
template <> struct AccOf<float> { typedef AccF T; };

/* spline value and derivative (w.r.t. x) on interval k with b = x/h - k */
template <typename R, typename Q>
static __device__ __forceinline__ R sp_val(const Q &q, R b)
{
    return q.x + b * (q.y + b * (q.z + b * q.w));
}
template <typename R, typename Q>
static __device__ __forceinline__ R sp_der(const Q &q, R b, R inv_h)
{
    return (q.y + b * ((R)2 * q.z + (R)3 * b * q.w)) * inv_h;
}

/* FP64 keeps the library's exp and sqrt and true divisions.  FP32 (the mixed
   build, whose bar is 1e-5 of max|F|) uses the hardware's single instructions:
   v_sqrt_f32, v_rsq_f32 and v_rcp_f32 (about 1 ulp) and v_exp_f32 on x log2(e), whose
   relative error grows as |x| eps, below 1e-6 for the Morse exponents met
   within a cutoff (|x| < 10) -- the library sequences are several times
   longer, and the Morse pass is bound by them. */
static __device__ __forceinline__ double fexp(double x) { return exp(x); }
static __device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
static __device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
static __device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
static __device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
static __device__ __forceinline__ double ffract(double x) { return __builtin_amdgcn_fract(x); }

/* tile position of scan trip g of stripe s: trips and stripes stay below 2^24, so the
   full-rate 24-bit multiply is exact (a 32-bit multiply issues at a quarter of the rate) */
static __device__ __forceinline__ int tpos(int s, int g, int S) { return s + (int)__umul24((unsigned)g, (unsigned)S); }
static __device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

/* LJ (lj.h:39-57) and Morse (morse.h:42-56) for one pair at r^2 inside the
   cutoff: the force on atom i is fac * (x_i - x_j), the energy share en */
template <typename R>
static __device__ __forceinline__ void lj_terms(const PairParam &p, R r2, R &fac, R &en)
{
    const R inv_r2 = frcp(r2);
    const R inv_rs2 = (R)p.p1 * inv_r2;
    const R inv_rs6 = inv_rs2 * inv_rs2 * inv_rs2;
    const R inv_rs12 = inv_rs6 * inv_rs6;
    fac = (R)12 * (R)p.p0 * inv_r2 * (inv_rs12 - (R)0.5 * inv_rs6);
    en = (R)0.5 * (R)p.p0 * (inv_rs12 - inv_rs6);
}
template <typename R>
static __device__ __forceinline__ void morse_terms(const PairParam &p, R r2, R &fac, R &en)
{
    if constexpr (sizeof(R) == sizeof(double))
    {
        const R r = fsqrt(r2);
        const R e1 = fexp(-(R)p.p1 * (r - (R)p.p2));
        const R e2 = e1 * e1;
        fac = (R)2 * (R)p.p1 * (R)p.p0 / r * (e2 - e1);
        en = (R)0.5 * (R)p.p0 * (e2 - (R)2 * e1);
    }
    else
    {
        /* one v_rsq_f32 gives both 1/r and r = r^2 / r; the exponent
           -alpha (r - r0) log2(e) is one FMA with per-pair-type constants */
        const R ir = __builtin_amdgcn_rsqf(r2);
        const R r = r2 * ir;
        const R ka = -(R)p.p1 * 1.44269504088896341f, kb = (R)p.p1 * (R)p.p2 * 1.44269504088896341f;
        const R e1 = __builtin_amdgcn_exp2f(__builtin_fmaf(ka, r, kb));
        const R e2 = e1 * e1;
        fac = (R)2 * (R)p.p1 * (R)p.p0 * ir * (e2 - e1);
        en = (R)0.5 * (R)p.p0 * (e2 - (R)2 * e1);
    }
}

/* one candidate pair inside the search radius; acc = {fx, fy, fz, energy or density} */
template <typename R, int PASS, int CAT, typename A, bool EN = true>
static __device__ __forceinline__ void pair_hit(const KDesc &D, const PairParam &p, R Fpi, R Fpj, R dx, R dy,
                                                R dz, A *acc)
{
    typedef typename Prec<R>::Q Q;
    const R r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < (R)p.rc2)) return;
    const int cat = CAT == CAT_GEN ? p.cat
                  : CAT == CAT_EAM1 ? FMDX_POT_EAM
                  : CAT == CAT_LJ ? FMDX_POT_LJ : FMDX_POT_MORSE;
    const EamDev &E = D.eam;
    if (PASS == 1)
    {
        if (cat != FMDX_POT_EAM) return;
        const R t = r2 * (R)E.inv_dr2;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        const R b = t - (R)k;
        const Q q = Prec<R>::rho(E)[(long)p.jloc * (E.nr2 - 1) + k];
        acc[3].add(sp_val<R, Q>(q, b));
        return;
    }
    if (cat == FMDX_POT_EAM)
    {
        const R inv = (R)E.inv_dr2;
        const R t = r2 * inv;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        const R b = t - (R)k;
        const long m = E.nr2 - 1;
        const Q *rt = Prec<R>::rho(E);
        const R rho_ip = sp_der<R, Q>(rt[(long)p.iloc * m + k], b, inv);
        const R rho_jp = (p.jloc == p.iloc) ? rho_ip : sp_der<R, Q>(rt[(long)p.jloc * m + k], b, inv);
        const Q q = Prec<R>::phi(E)[(long)(p.iloc * E.ne + p.jloc) * m + k];
        const R phid = sp_der<R, Q>(q, b, inv);
        const R mag = (R)2 * (Fpi * rho_jp + Fpj * rho_ip + phid);
        acc[0].add(-mag * dx);
        acc[1].add(-mag * dy);
        acc[2].add(-mag * dz);
        if (EN) acc[3].add((R)0.5 * sp_val<R, Q>(q, b));
    }
    else if (cat == FMDX_POT_LJ || cat == FMDX_POT_MORSE)
    {
        R fac, en;
        if (cat == FMDX_POT_LJ) lj_terms<R>(p, r2, fac, en);
        else morse_terms<R>(p, r2, fac, en);
        acc[0].add(fac * dx);
        acc[1].add(fac * dy);
        acc[2].add(fac * dz);
        if (EN) acc[3].add(en);
    }
}

/* ---- exact pairs of the FP32 kernels ----------------------------------------
   The FP32 separation of two atoms is off by up to ~1e-6 of rc^2 in r^2 (the
   rounding of the cell-relative positions, their image shift and the
   difference).  That decides nothing for a smooth pair function, but two kinds
   of pairs need the FP64 separation: a pair within that rounding of the cutoff
   is in or out of the reference's pair set (r^2 < rc^2, eam.h, lj.h, morse.h),
   and a tabulated function that is not smooth at the cutoff -- the setfl
   tables end in the two zero knots the reference appends (eam.c:444-471),
   which the natural spline turns into large oscillating curvature over the
   last intervals below rc^2 -- turns the rounding into large errors of the
   derivative there.  Pairs whose FP32 r^2 reaches rx2 (the lower of
   rc^2 (1 - 1e-5) and the start of the table's steep run, EamDev.rx2) are
   evaluated here in FP64 from the FP64 positions: the reference's pair set
   and the FP64 build's arithmetic, each pair term then rounded once to FP32
   into the compensated sums.  They are a thin shell, well under 1 % of the
   pairs. */
#define RX_BAND 1e-5   /* the rounding band below rc^2, relative */

struct ExactCtx
{
    const long *nb_base;   /* first slot of each of the 27 neighbour cells */
    const int *nb_pre;     /* their candidate-list offsets (28 entries) */
    const signed char (*w)[27];   /* their wraps */
    long is;               /* the home atom's slot */
};

/* FP64 separation x_i - x_j of the home atom and the candidate at place g of
   the 27-cell list, formed as the FP64 kernels form it (x_j + w L, then the
   difference); returns the candidate's slot */
static __device__ __forceinline__ long exact_sep(const KDesc &D, const ExactCtx &X, int g, double &dx, double &dy,
                                                 double &dz)
{
#pragma clang fp contract(off)
    int lo = 0, hi = 27;
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (X.nb_pre[mid] <= g) lo = mid;
        else hi = mid;
    }
    const long js = X.nb_base[lo] + (g - X.nb_pre[lo]);
    const double xj = D.x[js] + (double)X.w[0][lo] * D.L0;
    const double yj = D.y[js] + (double)X.w[1][lo] * D.L1;
    const double zj = D.z[js] + (double)X.w[2][lo] * D.L2;
    dx = D.x[X.is] - xj;
    dy = D.y[X.is] - yj;
    dz = D.z[X.is] - zj;
    return js;
}

/* one exact single-element EAM pair (eam.h:270-310 for the density, eam.h:177-228 for the force) */
template <int PASS, bool EN, typename A>
static __device__ __forceinline__ void exact_eam1(const KDesc &D, const ExactCtx &X, int g, double rc2, A *acc)
{
    double dx, dy, dz;
    const long js = exact_sep(D, X, g, dx, dy, dz);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < rc2)) return;
    const EamDev &E = D.eam;
    const double t = r2 * E.inv_dr2;
    int k = (int)t;
    if (k > E.nr2 - 2) k = E.nr2 - 2;
    const double b = t - (double)k;
    const double4 qr = E.rhoc[k];
    if (PASS == 1)
    {
        acc[3].add((float)sp_val<double, double4>(qr, b));
        return;
    }
    const double4 qp = E.phic[k];
    const double rho = sp_der<double, double4>(qr, b, E.inv_dr2);
    const double ph = sp_der<double, double4>(qp, b, E.inv_dr2);
    const double mag = 2.0 * (D.aux[X.is] * rho + D.aux[js] * rho + ph);
    acc[0].add((float)(-mag * dx));
    acc[1].add((float)(-mag * dy));
    acc[2].add((float)(-mag * dz));
    if (EN) acc[3].add((float)(0.5 * sp_val<double, double4>(qp, b)));
}

/* Coefficient records through a buffer resource: a lane without a pair passes
   an offset beyond the table, which the buffer unit answers with zeros and no
   memory access, so the gathers stay branch-free while only the lanes with a
   pair fetch.  The 128-bit results are reinterpreted as whole vectors:
   converting their elements one by one (bit_cast of each lane of the returned
   vector) has been compiled into a single 32-bit load whose value fills every
   component. */
typedef unsigned U4 __attribute__((ext_vector_type(4)));
typedef unsigned U2 __attribute__((ext_vector_type(2)));
typedef float F4 __attribute__((ext_vector_type(4)));
typedef float F2 __attribute__((ext_vector_type(2)));
typedef double D2 __attribute__((ext_vector_type(2)));
static __device__ __forceinline__ __amdgpu_buffer_rsrc_t rp_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
static __device__ __forceinline__ D2 ld2d(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
static __device__ __forceinline__ F4 ld4f(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(F4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
static __device__ __forceinline__ F2 ld2f(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(F2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
/* one interval of the density table (c0..c3) */
static __device__ __forceinline__ double4 rho_at(__amdgpu_buffer_rsrc_t r, unsigned off, double4 *)
{
    const D2 a = ld2d(r, off), b = ld2d(r, off + 16u);
    return make_double4(a.x, a.y, b.x, b.y);
}
static __device__ __forceinline__ float4 rho_at(__amdgpu_buffer_rsrc_t r, unsigned off, float4 *)
{
    const F4 a = ld4f(r, off);
    return make_float4(a.x, a.y, a.z, a.w);
}
#define RP_NONE 0x80000000u   /* an offset beyond any table */

/* Force record of one interval of a single-element table: the derivative
   coefficients of rho and phi first, scaled so that d/dx = r1 + b (r2 + b r3)
   (r1 = c1 / h, r2 = 2 c2 / h, r3 = 3 c3 / h, b = x / h - k), the values c0
   last, so that a force pass without energy reads 48 of the 64 bytes (FP64:
   three 16-byte loads instead of four) and the FP32 pass 24 of 32 (a 16- and
   an 8-byte load); the energy recovers c1..c3 from the scaled coefficients:
     FP64  (rho1 rho2)(rho3 phi1)(phi2 phi3)(rho0 phi0)
     FP32  (rho1 rho2 rho3 phi1)(phi2 phi3 rho0 phi0)            */
template <typename R> struct FRec { R r1, r2, r3, p1, p2, p3, p0; };
/* half the pair potential phi at position b of its interval, from the scaled coefficients:
   phi = c0 + b (c1 + b (c2 + b c3)) with c1 = h p1, c2 = h p2 / 2, c3 = h p3 / 3 */
template <typename R>
static __device__ __forceinline__ R pair_value(const FRec<R> &c, R b, R h)
{
    return (R)0.5 * (c.p0 + b * h * (c.p1 + b * ((R)0.5 * c.p2 + b * ((R)(1.0 / 3.0) * c.p3))));
}
template <bool EN>
static __device__ __forceinline__ FRec<double> frec_at(__amdgpu_buffer_rsrc_t r, unsigned off, double *)
{
    const D2 a = ld2d(r, off), b = ld2d(r, off + 16u), c = ld2d(r, off + 32u);
    FRec<double> q = {a.x, a.y, b.x, b.y, c.x, c.y, 0.0};
    if (EN) q.p0 = ld2d(r, off + 48u).y;
    return q;
}
template <bool EN>
static __device__ __forceinline__ FRec<float> frec_at(__amdgpu_buffer_rsrc_t r, unsigned off, float *)
{
    const F4 a = ld4f(r, off);
    FRec<float> q = {a.x, a.y, a.z, a.w, 0.0f, 0.0f, 0.0f};
    if (EN)
    {
        const F4 b = ld4f(r, off + 16u);
        q.p2 = b.x; q.p3 = b.y; q.p0 = b.w;
    }
    else
    {
        const F2 b = ld2f(r, off + 16u);
        q.p2 = b.x; q.p3 = b.y;
    }
    return q;
}

/* Single-element EAM: the set bits of a hit word, NP per iteration and
   without branches, so that the coefficient gathers of NP pairs are in
   flight together.  A slot without a pair (or outside the pair cutoff)
   reads zeros (no memory access) and contributes an exact zero.  Each
   accumulator takes the pairs in bit order whatever NP is.  The passes wait
   on the gathers' latency, so more pairs per iteration pay until the
   registers cost a wave per SIMD: measured on the 4M-atom copper step, FP64
   keeps 2 in both passes (3 in the force pass drops to 3 waves per SIMD and
   is 8 % slower; 3 or 4 held at 4 waves, by re-reading the tile after the
   gathers, gain nothing: the FP64 passes are bound by the gathers'
   throughput), FP32 takes 3 in the force pass and 4 in the density pass
   (FMDX_EAM_NP*, fmdx_tuned.h).  The density pass reads element 0's density
   table (32 or 16 bytes per interval), the force pass its force records; EN
   selects whether the force pass also takes the pair energy. */
template <typename R, int PASS, bool EN, bool LT, typename TileT, typename A>
static __device__ __forceinline__ void flush_eam1(const KDesc &D, const TileT &T, unsigned long long word, int s,
                                                  int g0, int S, R xi, R yi, R zi, R Fpi, double rc2,
                                                  unsigned long long &xw, A *acc, const float4 *lt)
{
    constexpr bool EX = sizeof(R) == sizeof(float);   /* FP32: pairs of the exact shell go to xw */
    const R rx2 = EX ? (R)D.eam.rx2 : (R)rc2;
    typedef typename Prec<R>::Q Q;
    constexpr int NP = sizeof(R) == sizeof(double) ? (PASS == 0 ? FMDX_EAM_NP64_0 : FMDX_EAM_NP64_1)
                                                   : (PASS == 0 ? FMDX_EAM_NP32_0 : FMDX_EAM_NP32_1);
    const EamDev &E = D.eam;
    const R inv = (R)E.inv_dr2;
    const R hh = (R)E.dr2;
    const int kmax = E.nr2 - 2;
    /* bytes per interval: density record (Q), force record (2 Q) */
    const unsigned rec = PASS == 1 ? (unsigned)sizeof(Q) : 2u * (unsigned)sizeof(Q);
    const __amdgpu_buffer_rsrc_t rb = rp_rsrc(PASS == 1 ? (const void *)Prec<R>::rho(E) : Prec<R>::rf(E),
                                              (unsigned)(E.nr2 - 1) * rec);
    const int m0 = tpos(s, g0, S);   /* the lane's entry of the word's first trip */
    const unsigned st = T.stride(S);
    while (__ballot(word != 0ull))
    {
        R dx[NP], dy[NP], dz[NP], fj[NP], b[NP];
        bool u[NP];
        unsigned o[NP];
        int kk[NP];
#pragma unroll
        for (int j = 0; j < NP; j++)
        {
            const bool v = word != 0ull;
            const int t = v ? __builtin_ctzll(word) : 0;
            word &= word - 1ull;
            R x, y, z;
            T.load(m0, st, t, x, y, z, fj[j]);
            dx[j] = xi - x;
            dy[j] = yi - y;
            dz[j] = zi - z;
            R r2 = dx[j] * dx[j] + dy[j] * dy[j] + dz[j] * dz[j];
            u[j] = v && r2 < rx2;
            if (EX) xw |= (v && r2 >= rx2) ? 1ull << t : 0ull;
            r2 = u[j] ? r2 : (R)0;
            /* interval k and the position b in it; r2 < rc^2 keeps k at most kmax + 1, whose
               offset lies beyond the table and reads zeros (the table's appended zero knots) */
            const R tt = r2 * inv;
            const int k = (int)tt;
            kk[j] = min(k, kmax);
            /* FP64: fract(tt), one operation; FP32: r2 inv - k from the unrounded product, the more
               accurate of the two where tt carries only 24 bits; the LDS table has no interval
               beyond kmax */
            if constexpr (LT) b[j] = tt - (R)kk[j];
            else if constexpr (sizeof(R) == sizeof(double)) b[j] = ffract(tt);
            else b[j] = __builtin_fmaf(r2, inv, -(float)k);
            o[j] = u[j] ? (unsigned)k * rec : RP_NONE;
        }
        if constexpr (LT)
        {
            /* the table in LDS: density coefficients (c0..c3) per interval, or the derivative
               knots (rho D, phi D, rho Y, phi Y) with dy/db = D_k + b (6 Y_k + 3 b (Y_k+1 - Y_k));
               a lane without a pair reads interval 0 and adds an exact zero */
            if (PASS == 1)
            {
#pragma unroll
                for (int j = 0; j < NP; j++)
                {
                    const R c = sp_val<R, float4>(lt[kk[j]], b[j]);
                    acc[3].add(u[j] ? c : (R)0);
                }
            }
            else
            {
                float4 k0[NP];
                float2 y1[NP];
#pragma unroll
                for (int j = 0; j < NP; j++)
                {
                    k0[j] = lt[kk[j]];
                    y1[j] = *reinterpret_cast<const float2 *>(&lt[kk[j] + 1].z);
                }
#pragma unroll
                for (int j = 0; j < NP; j++)
                {
                    const R bb = b[j];
                    const R rho = (k0[j].x + bb * ((R)6 * k0[j].z + (R)3 * bb * (y1[j].x - k0[j].z))) * inv;
                    const R ph = (k0[j].y + bb * ((R)6 * k0[j].w + (R)3 * bb * (y1[j].y - k0[j].w))) * inv;
                    const R mag = (R)2 * (Fpi * rho + fj[j] * rho + ph);
                    acc[0].add(u[j] ? -mag * dx[j] : (R)0);
                    acc[1].add(u[j] ? -mag * dy[j] : (R)0);
                    acc[2].add(u[j] ? -mag * dz[j] : (R)0);
                    if (EN)
                    {
                        /* the pair energy from phi's value coefficients (c0..c3) in memory: one
                           16-byte load, half the force record's offset */
                        const __amdgpu_buffer_rsrc_t rv = rp_rsrc(Prec<R>::phi(E), (unsigned)(E.nr2 - 1) * 16u);
                        const float4 qv = rho_at(rv, u[j] ? o[j] / 2u : RP_NONE, (float4 *)nullptr);
                        const R e = (R)0.5 * sp_val<R, float4>(qv, bb);
                        acc[3].add(u[j] ? e : (R)0);
                    }
                }
            }
        }
        else if (PASS == 1)
        {
            Q q[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) q[j] = rho_at(rb, o[j], (Q *)nullptr);
            R part = (R)0;   /* FP32: the iteration's pair terms summed before one compensated add */
#pragma unroll
            for (int j = 0; j < NP; j++)
            {
                const R c = sp_val<R, Q>(q[j], b[j]);
                /* FP32: an empty slot read a zero record at b = 0, so that c = 0 */
                if constexpr (EX) part += c;
                else acc[3].add(u[j] ? c : (R)0);
            }
            if constexpr (EX) acc[3].add(part);
        }
        else
        {
            FRec<R> q[NP];
            R pxs = (R)0, pys = (R)0, pzs = (R)0, pes = (R)0;
#pragma unroll
            for (int j = 0; j < NP; j++) q[j] = frec_at<EN>(rb, o[j], (R *)nullptr);
#pragma unroll
            for (int j = 0; j < NP; j++)
            {
                const FRec<R> &c = q[j];
                const R bb = b[j];
                const R rho = c.r1 + bb * (c.r2 + bb * c.r3);
                const R ph = c.p1 + bb * (c.p2 + bb * c.p3);
                const R mag = (R)2 * (Fpi * rho + fj[j] * rho + ph);
                const R e = EN ? pair_value(c, bb, hh) : (R)0;
                if constexpr (EX)
                {
                    /* an empty slot read a zero record: mag = 0 and e = 0, times a finite
                       separation (tile entries and padding are finite) */
                    pxs = __builtin_fmaf(-mag, dx[j], pxs);
                    pys = __builtin_fmaf(-mag, dy[j], pys);
                    pzs = __builtin_fmaf(-mag, dz[j], pzs);
                    if (EN) pes += e;
                }
                else
                {
                    acc[0].add(u[j] ? -mag * dx[j] : (R)0);
                    acc[1].add(u[j] ? -mag * dy[j] : (R)0);
                    acc[2].add(u[j] ? -mag * dz[j] : (R)0);
                    if (EN) acc[3].add(u[j] ? e : (R)0);
                }
            }
            if constexpr (EX)
            {
                acc[0].add(pxs);
                acc[1].add(pys);
                acc[2].add(pzs);
                if (EN) acc[3].add(pes);
            }
        }
    }
}

/* Single-kind LJ or Morse: the set bits of a hit word two per iteration
   without branches, as flush_eam1 does; a slot without a pair evaluates the
   formulas at r^2 = rc^2 and adds exact zeros. */
/* Each pair once (FMDX_N3, FP32 single-kind LJ and Morse): the home cell stages itself
   and its 13 forward neighbours; a pair with a candidate of a forward cell adds the
   force on the home atom to that atom's sums and its opposite to the candidate's entry
   of the tile's partner sums (LDS, one wave: a fixed order), and its whole energy to
   the home atom.  Pairs within the home cell are met from both sides, as before: home
   atom's force only, half the energy each.  Tile entries below jmin are home-cell
   candidates. */
struct JSide
{
    float *jx, *jy, *jz;
    int jmin;
};

/* forward direction f (0..12): the offsets after (0, 0, 0) in (x, y, z) order,
   (1, *, *), (0, 1, *), (0, 0, 1) */
static __device__ __forceinline__ void n3_offset(int f, int &ox, int &oy, int &oz)
{
    ox = f < 9 ? 1 : 0;
    oy = f < 9 ? f / 3 - 1 : (f < 12 ? 1 : 0);
    oz = f < 9 ? f % 3 - 1 : (f < 12 ? f - 10 : 1);
}

/* one exact single-kind LJ or Morse pair (lj.h:26-58, morse.h:26-58); N3: m is the
   candidate's tile entry */
template <int CAT, bool EN, bool N3, typename A>
static __device__ __forceinline__ void exact_pair1(const KDesc &D, const ExactCtx &X, int g, const PairParam &p, A *acc,
                                                   int m, const JSide &J)
{
    double dx, dy, dz;
    exact_sep(D, X, g, dx, dy, dz);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < p.rc2)) return;
    double fac, en;
    if (CAT == CAT_LJ) lj_terms<double>(p, r2, fac, en);
    else morse_terms<double>(p, r2, fac, en);
    const float fx = (float)(fac * dx), fy = (float)(fac * dy), fz = (float)(fac * dz);
    acc[0].add(fx);
    acc[1].add(fy);
    acc[2].add(fz);
    const bool fw = N3 && m >= J.jmin;
    if constexpr (N3)
    {
        if (fw)
        {
            atomicAdd(&J.jx[m], -fx);
            atomicAdd(&J.jy[m], -fy);
            atomicAdd(&J.jz[m], -fz);
        }
    }
    if (EN) acc[3].add(fw ? (float)(2.0 * en) : (float)en);
}

/* one exact pair of any category (the generic dispatch of pair_hit in FP64) */
template <int PASS, bool EN, typename A>
static __device__ __forceinline__ void exact_pair(const KDesc &D, const ExactCtx &X, int g, const PairParam &p, A *acc)
{
    double dx, dy, dz;
    const long js = exact_sep(D, X, g, dx, dy, dz);
    AccD t[4];
    for (int q = 0; q < 4; q++) t[q].zero();
    const double Fpi = (PASS == 0 && D.has_eam) ? D.aux[X.is] : 0.0;
    const double Fpj = (PASS == 0 && D.has_eam) ? D.aux[js] : 0.0;
    pair_hit<double, PASS, CAT_GEN, AccD, EN>(D, p, Fpi, Fpj, dx, dy, dz, t);
    for (int q = 0; q < 4; q++) acc[q].add((float)t[q].get());
}

template <typename R, int CAT, bool EN, bool N3, typename TileT, typename A>
static __device__ __forceinline__ void flush_pair2(const KDesc &D, const TileT &T, unsigned long long word, int s,
                                                   int g0, int S, R xi, R yi, R zi, const PairParam &p,
                                                   bool &ex, A *acc, const JSide &J)
{
    constexpr bool EX = sizeof(R) == sizeof(float);   /* FP32: a pair in the rounding band below rc^2 sets ex */
    const R rc2 = (R)p.rc2;
    const R rx2 = EX ? (R)(p.rc2 * (1.0 - RX_BAND)) : rc2;
    /* FP32: the pair terms of the whole word are summed in plain FP32 and compensated
       once, after the word (batched LJ/Morse ensemble 2.73 -> 2.63 ms per step against one
       compensated add per iteration, same max mixed/FP64 force error 7.1e-6) */
    R wpx = (R)0, wpy = (R)0, wpz = (R)0, wpe = (R)0;
    const int m0 = tpos(s, g0, S);   /* the lane's entry of the word's first trip */
    const unsigned st = T.stride(S);
    if (__ballot(word != 0ull)) do
    {
        const bool v1 = word != 0ull;
        const int t1 = v1 ? __builtin_ctzll(word) : 0;
        word &= word - 1ull;
        const bool v2 = word != 0ull;
        const int t2 = v2 ? __builtin_ctzll(word) : 0;
        word &= word - 1ull;
        R x1, y1, z1, f1, x2, y2, z2, f2;
        T.load(m0, st, t1, x1, y1, z1, f1);
        T.load(m0, st, t2, x2, y2, z2, f2);
        const R dx1 = xi - x1, dy1 = yi - y1, dz1 = zi - z1;
        const R dx2 = xi - x2, dy2 = yi - y2, dz2 = zi - z2;
        R r21 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
        R r22 = dx2 * dx2 + dy2 * dy2 + dz2 * dz2;
        const bool in1 = r21 < rx2, in2 = r22 < rx2;
        const bool u1 = v1 && in1, u2 = v2 && in2;
        /* the separations of a live lane's slots are finite: not below rx2 is at least rx2 */
        if (EX) ex = ex || (v1 && !in1) || (v2 && !in2);
        r21 = u1 ? r21 : rc2;
        r22 = u2 ? r22 : rc2;
        R fac1, en1, fac2, en2;
        if (CAT == CAT_LJ) { lj_terms<R>(p, r21, fac1, en1); lj_terms<R>(p, r22, fac2, en2); }
        else { morse_terms<R>(p, r21, fac1, en1); morse_terms<R>(p, r22, fac2, en2); }
        if constexpr (EX)
        {
            /* an empty slot's factor is zeroed and multiplies a finite separation (tile
               entries, padding included, and home coordinates are finite) */
            const R f1 = u1 ? fac1 : (R)0, f2 = u2 ? fac2 : (R)0;
            wpx = __builtin_fmaf(f2, dx2, __builtin_fmaf(f1, dx1, wpx));
            wpy = __builtin_fmaf(f2, dy2, __builtin_fmaf(f1, dy1, wpy));
            wpz = __builtin_fmaf(f2, dz2, __builtin_fmaf(f1, dz1, wpz));
            if constexpr (N3)
            {
                /* the forward candidates' opposite forces, only from lanes with such a pair: the
                   empty slots of a stripe all name one entry, and adds of theirs to it would
                   serialise in the LDS */
                const int m1 = m0 + (int)__umul24((unsigned)t1, (unsigned)S), m2 = m0 + (int)__umul24((unsigned)t2, (unsigned)S);
                const bool w1 = u1 && m1 >= J.jmin, w2 = u2 && m2 >= J.jmin;
                if (w1)
                {
                    atomicAdd(&J.jx[m1], -f1 * dx1);
                    atomicAdd(&J.jy[m1], -f1 * dy1);
                    atomicAdd(&J.jz[m1], -f1 * dz1);
                }
                if (w2)
                {
                    atomicAdd(&J.jx[m2], -f2 * dx2);
                    atomicAdd(&J.jy[m2], -f2 * dy2);
                    atomicAdd(&J.jz[m2], -f2 * dz2);
                }
                if (EN) wpe += (u1 ? (w1 ? (R)2 * en1 : en1) : (R)0) + (u2 ? (w2 ? (R)2 * en2 : en2) : (R)0);
            }
            else if (EN) wpe += (u1 ? en1 : (R)0) + (u2 ? en2 : (R)0);
        }
        else
        {
            acc[0].add(u1 ? fac1 * dx1 : (R)0);
            acc[1].add(u1 ? fac1 * dy1 : (R)0);
            acc[2].add(u1 ? fac1 * dz1 : (R)0);
            if (EN) acc[3].add(u1 ? en1 : (R)0);
            acc[0].add(u2 ? fac2 * dx2 : (R)0);
            acc[1].add(u2 ? fac2 * dy2 : (R)0);
            acc[2].add(u2 ? fac2 * dz2 : (R)0);
            if (EN) acc[3].add(u2 ? en2 : (R)0);
        }
    } while (__ballot(word != 0ull));
    if constexpr (EX)
    {
        acc[0].add(wpx);
        acc[1].add(wpy);
        acc[2].add(wpz);
        if (EN) acc[3].add(wpe);
    }
}

/* the bits of hit word w whose pairs have an FP32 r^2 of at least rx2 */
template <typename TileT>
static __device__ unsigned long long band_bits(const TileT &T, unsigned long long w, int s, int g0, int S, float xi,
                                               float yi, float zi, float rx2)
{
    unsigned long long out = 0ull;
    while (w)
    {
        const int t = __builtin_ctzll(w);
        w &= w - 1ull;
        float x, y, z, f;
        T.load(tpos(s, g0 + t, S), x, y, z, f);
        const float dx = xi - x, dy = yi - y, dz = zi - z;
        if (dx * dx + dy * dy + dz * dz >= rx2) out |= 1ull << t;
    }
    return out;
}

/* the pairs of the exact shell collected from one hit word (bit t: trip t of
   this lane's stripe), evaluated in FP64 after the word's FP32 flush */
template <int PASS, int CAT, bool EN, bool N3, typename TileT, typename A>
static __device__ __forceinline__ void flush_exact(const KDesc &D, const TileT &T, unsigned long long xw, int s, int g0, int S,
                                   const ExactCtx &X, const PairParam *ppi, const PairParam &p1, A *acc,
                                   const JSide &J)
{
    while (__ballot(xw != 0ull))
    {
        if (xw)
        {
            const int t = __builtin_ctzll(xw);
            xw &= xw - 1ull;
            const int m = tpos(s, g0 + t, S);
            if constexpr (CAT == CAT_EAM1) exact_eam1<PASS, EN>(D, X, T.g[m], p1.rc2, acc);
            else if constexpr (CAT == CAT_LJ || CAT == CAT_MORSE) exact_pair1<CAT, EN, N3>(D, X, T.g[m], p1, acc, m, J);
            else exact_pair<PASS, EN>(D, X, T.g[m], ppi[T.k[m]], acc);
        }
    }
}

/* EN: the force pass also sums the pair energy into cell_pe (the density pass always keeps the
   embedding energy of each cell, a by-product of F'(rho)) */
/* The lanes of one home cell meet.  In the LDS-table kernels each wave of the
   workgroup owns its own home cell, so the meeting is the wave's: one wave's
   LDS accesses complete in the order it issues them, and the fences keep the
   compiler from moving memory accesses across. */
template <bool LT>
static __device__ __forceinline__ void csync()
{
    if constexpr (LT)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    else
        __syncthreads();
}

/* offset of the LDS table after the pair parameters in the dynamic LDS of the LDS-table kernels */
static __host__ __device__ __forceinline__ unsigned lt_offset(int nkinds)
{
    return (unsigned)((sizeof(PairParam) * (size_t)nkinds * (size_t)nkinds + 15u) & ~(size_t)15u);
}

#if FMDX_EAM_LDS
#define LW FMDX_EAM_LDS   /* waves, each with its own home cell, per workgroup of the LDS-table kernels */
#else
#define LW 1
#endif

/* One home cell's pair sums.  LT: the FP32 single-element EAM passes with the
   spline table resident in LDS (k_cell_lds): the calling wave is one of LW
   in its workgroup, each with its own tile, and the table follows the pair
   parameters in the dynamic LDS. */
template <typename R, int PASS, int CAT, bool EN, bool LT = false, bool N3 = false>
static __device__ __forceinline__ void b_cell(const KDesc &D, const int c)
{
    typedef typename AccOf<R>::T Acc;
    typedef Prec<R> P;
    static_assert(!LT || (sizeof(R) == sizeof(float) && CAT == CAT_EAM1 && FMDX_WAVES_PER_CELL == 1),
                  "LDS tables: FP32 single-element EAM, one wave per cell");
    static_assert(!N3 || (sizeof(R) == sizeof(float) && (CAT == CAT_LJ || CAT == CAT_MORSE) && PASS == 0 && !LT &&
                          FMDX_WAVES_PER_CELL == 1),
                  "each pair once: FP32 single-kind LJ or Morse, one wave per cell");
    /* neighbour cells staged: all 27, or (N3) the home cell and its 13 forward neighbours;
       the list entry of the home cell */
    constexpr int NNB = N3 ? 14 : 27, HOME = N3 ? 0 : 13;
    /* candidates per tile, per precision and pair category (fmdx_tuned.h): the FP32 EAM kernels
       keep a smaller tile, so that the density pass's LDS (tile plus the list places of the exact
       pairs) still lets 20 workgroups share a CU */
    enum
    {
        CT = LT ? FMDX_EAM_LDS_TILE
           : sizeof(R) == sizeof(double) ? (int)P::CT
           : N3 ? FMDX_TILE32_N3
           : CAT == CAT_EAM1 ? FMDX_TILE32_EAM : CAT == CAT_LJ ? FMDX_TILE32_LJ
           : CAT == CAT_MORSE ? FMDX_TILE32_MORSE : (int)P::CT
    };
    constexpr bool WITH_FP = PASS == 0 && (CAT == CAT_GEN || CAT == CAT_EAM1);
    constexpr int NWS = LT ? LW : 1;                 /* home cells sharing the workgroup's LDS */
    const int lane = LT ? (int)(threadIdx.x & (WG - 1)) : (int)threadIdx.x;   /* thread within the home cell */
    const int ws = LT ? (int)(threadIdx.x / WG) : 0;  /* its home cell's LDS slot */
    if ((PASS == 1 && !D.has_eam) || D.cnt[c] == 0)
    {
        if (lane == 0)
        {
            if (PASS == 1) { D.cell_emb[c] = 0.0; D.htrips[c] = -1; }
            else if (EN) D.cell_pe[c] = 0.0;
        }
        return;
    }
    __shared__ long nb_base_s[NWS][27];
    __shared__ __align__(16) int nb_pre_s[NWS][32];   /* 28 prefixes, then INT_MAX */
    __shared__ Shifts<R> SH_s[NWS];
    __shared__ Tile<R, WITH_FP, CAT == CAT_GEN, CT> T_s[NWS];
    /* N3: the partner sums of the tile's entries (x, y, z) */
    __shared__ float jside_s[N3 ? 3 : 1][N3 ? CT + NT : 1];
    long *const nb_base = nb_base_s[ws];
    int *const nb_pre = nb_pre_s[ws];
    Shifts<R> &SH = SH_s[ws];
    auto &T = T_s[ws];

    extern __shared__ PairParam pp[];   /* nkinds^2 entries, sized at launch */
    const float4 *const lt = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(pp) + lt_offset(D.nkinds));
    /* the stripe sums reuse the tile once the scan is over */
    double(*red)[NT] = reinterpret_cast<double(*)[NT]>(&T);
    static_assert(sizeof(T) >= 4 * NT * sizeof(double), "tile too small for the stripe sums");
    /* with several waves per home cell: each wave's survivors of a staging batch (double-buffered
       by batch parity), and the tile position of each home atom of the block */
    __shared__ int wcnt[2][FMDX_WAVES_PER_CELL];
    __shared__ short selfpos[FMDX_WAVES_PER_CELL > 1 ? NT : 1];

    const int wv = lane / WG;       /* the wave within the home cell */
    const int nkk = D.nkinds * D.nkinds;
    if (!LT)   /* the LDS-table kernels load them once */
        for (int i = lane; i < nkk; i += NT) pp[i] = D.pairs[i];
    const int cz = c % D.nc2, ct = c / D.nc2, cy = ct % D.nc1, cx = ct / D.nc1;
    /* the density pass's record of this cell, loaded here so that it arrives with the neighbour counts */
    int hw_ld = -1;
    unsigned long long tiles_ld = 0ull;
    if (PASS == 0 && D.has_eam)
    {
        hw_ld = D.htrips[c];
        tiles_ld = D.htile[c];
    }
    int my_cnt = 0;
    if (lane < NNB)
    {
        int ox = lane / 9 - 1, oy = (lane / 3) % 3 - 1, oz = lane % 3 - 1;
        if constexpr (N3)
        {
            /* entry 0 the home cell, entry 1 + f forward direction f */
            if (lane == 0) ox = oy = oz = 0;
            else n3_offset(lane - 1, ox, oy, oz);
        }
        int jx = cx + ox, jy = cy + oy, jz = cz + oz;
        int wx = 0, wy = 0, wz = 0;
        bool ok = true;
        if (jx < 0) { if (D.pbc0) { jx += D.nc0; wx = -1; } else ok = false; }
        else if (jx >= D.nc0) { if (D.pbc0) { jx -= D.nc0; wx = 1; } else ok = false; }
        if (jy < 0) { if (D.pbc1) { jy += D.nc1; wy = -1; } else ok = false; }
        else if (jy >= D.nc1) { if (D.pbc1) { jy -= D.nc1; wy = 1; } else ok = false; }
        if (jz < 0) { if (D.pbc2) { jz += D.nc2; wz = -1; } else ok = false; }
        else if (jz >= D.nc2) { if (D.pbc2) { jz -= D.nc2; wz = 1; } else ok = false; }
        const long jc = ok ? ((long)jx * D.nc1 + jy) * D.nc2 + jz : 0;
        nb_base[lane] = jc * D.cap;
        my_cnt = ok ? D.cnt[jc] : 0;
        const int wr[3] = {wx, wy, wz}, dj[3] = {jx, jy, jz}, hc[3] = {cx, cy, cz};
        SH.set(D, lane, wr, dj, hc);
    }
    /* inclusive prefix of the neighbour counts over lanes 0..26 (lanes 27..31 hold 0):
       within each row of 16 lanes by DPP row shifts (a lane whose source lies outside
       its row adds 0), then lanes 16..31 add lane 15's total */
    int incl = my_cnt;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, false);   /* row_shr:1 */
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, false);   /* row_shr:2 */
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, false);   /* row_shr:4 */
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, false);   /* row_shr:8 */
    {
        const int row0 = __builtin_amdgcn_readlane(incl, 15);
        if ((lane & (WG - 1)) >= 16 && (lane & (WG - 1)) < 32) incl += row0;
    }
    if (lane < NNB) nb_pre[lane + 1] = incl;
    else if (lane < 31) nb_pre[lane + 1] = 0x7fffffff;
    if (lane == 0) nb_pre[0] = 0;
    csync<LT>();
    const int ncand = __builtin_amdgcn_readfirstlane(nb_pre[NNB]);
    /* Neighbour cell of candidate g of the 27-cell list: the last prefix not above g
       (empty cells share their successor's prefix and are passed over).  Two levels:
       the block of eight prefixes from three wave-uniform pivots, then one 32-byte
       read of that block: one LDS round trip per staging batch where a binary search
       over the 27 prefixes takes five dependent ones (4M-atom copper step 3.50 -> 3.43
       ms FP64, 2.91 -> 2.80 ms mixed). */
    const int pv8 = __builtin_amdgcn_readfirstlane(nb_pre[8]);
    const int pv16 = __builtin_amdgcn_readfirstlane(nb_pre[16]);
    const int pv24 = __builtin_amdgcn_readfirstlane(nb_pre[24]);
    auto locate = [&](int g) -> int {
        const int blk = 8 * ((g >= pv8) + (g >= pv16) + (g >= pv24));
        const int4 u = *reinterpret_cast<const int4 *>(nb_pre + blk);
        const int4 v = *reinterpret_cast<const int4 *>(nb_pre + blk + 4);
        return blk + (u.y <= g) + (u.z <= g) + (u.w <= g) + (v.x <= g) + (v.y <= g) + (v.z <= g) + (v.w <= g);
    };

    const int nh = __builtin_amdgcn_readfirstlane(D.cnt[c]);
    const int ag = D.ag;
    const R rcut2 = (R)D.rcmax2;
    /* FP32 tiles without F' keep |x_j|^2 in their fourth word, and the scan forms
       d = |x_j|^2 + (|x_i|^2 - rc^2 (1 + 2.5e-4)) - 2 x_i.x_j with one add and three
       FMAs and shifts the sign of d into the hit word (v_alignbit), where the
       difference form takes six operations, a compare and a select.  The scan only
       selects candidates: every selected pair is tested again with the differences
       (flush_eam1, pair_hit), so the margin only has to exceed the cancellation
       error; cell-relative coordinates stay below 4 sqrt(3) rc (cells are narrower
       than 2 rc), which bounds that error by ~2e-5 rc^2.  No NaN may reach the d of
       a live home atom, whose sign bit would be arbitrary: padding entries are
       (0, 0, 0, +inf), which gives d = +inf.  The other tiles keep the difference
       form, with the margin folded into the first FMA,
       d = dz^2 + (dy^2 + (dx^2 - rc^2 (1 + m))), and the same sign shift (m = 1e-12
       in FP64); there padding sits at (-1e30, -1e30, -1e30) (FP64: -1e300), whose
       squared distance overflows, so that d = +inf again.  A lane without
       a live home atom scans with zero multipliers (norm form) or from +inf
       (difference form), and its word is cleared whatever it read. */
    constexpr bool NSCAN = sizeof(R) == sizeof(float) && !WITH_FP;
    const R rscan2 = rcut2 * (R)(sizeof(R) == sizeof(double) ? 1.0 + 1e-12 : 1.0 + 2.5e-4);
    const R pinf = (R)__builtin_inf();
    /* padding of the difference-form tiles: finite, but so far that its squared
       distance from any home atom overflows to +inf */
    const R pfar = sizeof(R) == sizeof(float) ? (R)1e30f : (R)1e300;
    const long hbase = (long)c * D.cap;
    /* Home atoms per block.  One block holds up to one atom per lane, each atom with
       64 / nh stripes, so that a cell of 38 atoms leaves 26 lanes idle through the scan
       and the flush.  Cells of more than 32 atoms (up to FMDX_SPLIT_MAX) are taken in
       blocks of 32 atoms, two stripes each, and the rest, the tile staged for each
       block (38 atoms: nt / 2 + nt / 10 scan trips instead of nt). */
    const int bsz = (!N3 && FMDX_WAVES_PER_CELL == 1 && nh > WG / 2 && nh <= FMDX_SPLIT_MAX) ? WG / 2 : NT;
    /* N3: partner sums per forward direction d (0..12) and slot, x then y then z arrays */
    const long nslot = (long)D.ncell * D.cap;
    const bool one_block = nh > 0 && nh <= bsz;
    const int hw = __builtin_amdgcn_readfirstlane(hw_ld);
    /* a density pass without hit-word storage (hbits null) records nothing: the force pass rescans */
    const bool record = PASS == 1 && one_block && D.hbits != nullptr;
    if (PASS == 1 && lane == 0 && !record) D.htrips[c] = -1;
    bool replay = PASS == 0 && one_block && hw >= 0;
    /* tile sizes of the density pass's first four rounds (16 bits each): the force pass replays
       a round only if it staged as many candidates, and rescans from the first round that differs */
    unsigned long long tiles = replay ? tiles_ld : 0ull;
    PairParam p1;
    if (CAT != CAT_GEN) p1 = pp[0];
    /* the home cell's box in the frame of the staged positions, and the culling
       radius with a margin over the rounding of the box distance */
    R bx0, bx1, by0, by1, bz0, bz1;
    if (sizeof(R) == sizeof(double))
    {
        bx0 = (R)(cx * D.h0); bx1 = (R)((cx + 1) * D.h0);
        by0 = (R)(cy * D.h1); by1 = (R)((cy + 1) * D.h1);
        bz0 = (R)(cz * D.h2); bz1 = (R)((cz + 1) * D.h2);
    }
    else
    {
        bx0 = by0 = bz0 = (R)0;
        bx1 = (R)D.h0; by1 = (R)D.h1; bz1 = (R)D.h2;
    }
    const R cull2 = rcut2 * (sizeof(R) == sizeof(double) ? (R)(1.0 + 1e-12) : (R)(1.0 + 1e-5));
    double cell_acc = 0.0;

    for (int h0 = 0; h0 < nh; h0 += bsz)
    {
        const int nhc = min(bsz, nh - h0);
        const int S = NT / nhc;
        const int a = lane % nhc, s = lane / nhc;
        const long is = hbase + h0 + a;
        const ExactCtx X = {nb_base, nb_pre, SH.w, is};
        /* a lane without an active home atom (live false) scans so that it finds no
           candidate, from finite coordinates, so that the flushes may multiply its
           empty slots' zero factors by finite separations */
        R xi = 0, yi = 0, zi = 0, Fpi = 0;
        int ki = 0;
        bool act = false;
        if (s < S)
        {
            ki = D.kind[is];
            act = (ag < 0 || D.grp[is] == ag);
            if (act)
            {
                R unused;
                P::template load<false>(D, is, xi, yi, zi, unused);
            }
            if (PASS == 0 && D.has_eam) Fpi = (R)D.aux[is];
        }
        const bool live = act && s < S;
        const PairParam *ppi = pp + ki * D.nkinds;
        Acc acc[4];
        for (int q = 0; q < 4; q++) acc[q].zero();
        int wi = 0;   /* hit words so far (wave-uniform) */
        int gpos = 0; /* next candidate of the 27-cell list to stage (wave-uniform) */
        int round = 0;
        while (gpos < ncand)
        {
            csync<LT>();   /* the previous tile is no longer read */
            /* the round's first recorded hit word loads while the tile is staged */
            const unsigned long long wfirst =
                (replay && wi < hw) ? D.hbits[((long)c * HBITS_WORDS + wi) * NT + lane] : 0ull;
            /* fill the tile with the candidates that can reach the home cell: a
               candidate farther than the cutoff from the cell's box cannot pair
               with any home atom; survivors are compacted in list order */
            int nt = 0;
            int jmin = 0;   /* N3: home-cell candidates staged in this round (the tile's first entries) */
            /* tile position of this lane's home atom in this round, -1 while it is not in the tile */
            int selft = -1;
            const int gself = nb_pre[HOME] + h0 + a;   /* its place in the candidate list */
            if constexpr (FMDX_WAVES_PER_CELL > 1)
            {
                if (lane < nhc) selfpos[lane] = -1;
            }
            int batch = 0;
            /* one batch of candidates as loaded: position, F' (and kind, which only the
               generic kernel dispatches on) load together, whether or not the candidate
               survives the culling: a load issued after the test would add one more round
               trip to every batch (force pass 1.88 -> 1.80 ms FP64, 1.60 -> 1.52 ms mixed,
               on the 4M-atom copper step) */
            struct Cand
            {
                R x, y, z, f;
                int kj, lo, gr;
                long js;
            };
            auto fetch = [&](int g, Cand &q) {
                q.x = q.y = q.z = q.f = (R)0;
                q.kj = 0; q.lo = 0; q.gr = ag; q.js = 0;
                if (g < ncand)
                {
                    q.lo = locate(g);
                    const long js = nb_base[q.lo] + (g - nb_pre[q.lo]);
                    q.js = js;
                    P::template load<WITH_FP>(D, js, q.x, q.y, q.z, q.f);
                    if (CAT == CAT_GEN) q.kj = D.kind[js];
                    if (ag >= 0) q.gr = D.grp[js];
                }
            };
            while (gpos < ncand && nt <= (int)CT - NT)
            {
                const int g = gpos + lane;
                Cand cur;
                fetch(g, cur);
                bool keep = false;
                R x = 0, y = 0, z = 0;
                const R fpv = cur.f;
                const int kj = cur.kj, lo = cur.lo;
                if (g < ncand)
                {
                    /* staged position and culling test without FMA contraction, so that every
                       instantiation keeps the same candidates: the force pass replays the
                       density pass's hit words by tile position */
#pragma clang fp contract(off)
                    x = cur.x + SH.sx(D, lo);
                    y = cur.y + SH.sy(D, lo);
                    z = cur.z + SH.sz(D, lo);
                    const R ex = fmax(fmax(bx0 - x, x - bx1), (R)0);
                    const R ey = fmax(fmax(by0 - y, y - by1), (R)0);
                    const R ez = fmax(fmax(bz0 - z, z - bz1), (R)0);
                    keep = ex * ex + ey * ey + ez * ez <= cull2 && (ag < 0 || cur.gr == ag);
                }
                const unsigned long long bal = __ballot(keep);
                int off = nt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
                int kept = __popcll(bal);
                if constexpr (FMDX_WAVES_PER_CELL > 1)
                {
                    /* survivors are compacted in list order across the waves */
                    const int par = batch & 1;
                    if ((lane & (WG - 1)) == 0) wcnt[par][wv] = kept;
                    csync<LT>();
                    kept = 0;
                    for (int w = 0; w < FMDX_WAVES_PER_CELL; w++)
                    {
                        if (w < wv) off += wcnt[par][w];
                        kept += wcnt[par][w];
                    }
                }
                if (keep)
                {
                    T.put(off, x, y, z, NSCAN ? x * x + y * y + z * z : fpv, kj);
                    if constexpr (sizeof(R) == sizeof(float)) T.g[off] = (unsigned short)g;
                    if constexpr (N3)
                    {
                        jside_s[0][off] = 0.0f;
                        jside_s[1][off] = 0.0f;
                        jside_s[2][off] = 0.0f;
                    }
                }
                if constexpr (N3)
                {
                    /* home-cell candidates come first in list order; a forward candidate that
                       is not staged gets a zero partner sum from this cell (written once, by
                       the first block of home atoms) */
                    jmin += __popcll(__ballot(keep && lo == HOME));
                    if (h0 == 0 && g < ncand && !keep && lo != HOME)
                    {
                        float *const pj = D.np + (long)(lo - 1) * nslot + cur.js;
                        pj[0] = 0.0f;
                        pj[13 * nslot] = 0.0f;
                        pj[26 * nslot] = 0.0f;
                    }
                }
                /* the lane that staged this lane's home atom (list entry HOME is the home
                   cell itself, not an image) hands over the tile position it wrote */
                if constexpr (FMDX_WAVES_PER_CELL == 1)
                {
                    const int src = gself - gpos;
                    const int got = __shfl((keep && lo == HOME) ? off : -1, src & (WG - 1));
                    if (src >= 0 && src < WG) selft = got;
                }
                else
                {
                    const int hk = g - nb_pre[HOME] - h0;
                    if (keep && lo == HOME && hk >= 0 && hk < nhc) selfpos[hk] = (short)off;
                }
                nt += kept;
                gpos += NT;
                batch++;
            }
            {
                const unsigned sh16 = 16u * (unsigned)min(round, 3);
                if (record && round < 4) tiles |= (unsigned long long)nt << sh16;
                const unsigned rec = round < 4 ? (unsigned)((tiles >> sh16) & 0xFFFFull) : 0u;
                if (replay && rec != (unsigned)nt)
                {
                    replay = false;
                    if (lane == 0) atomicAdd(&D.counters[6], 1u);
                }
                round++;
            }
            const int trips = (nt + S - 1) / S;
            const int npad = trips * S;
            for (int m = nt + lane; m < npad; m += NT)
            {
                if constexpr (NSCAN) T.put(m, (R)0, (R)0, (R)0, pinf, NOKIND);
                else T.put(m, -pfar, -pfar, -pfar, (R)0, NOKIND);
            }
            csync<LT>();
            if constexpr (FMDX_WAVES_PER_CELL > 1) selft = selfpos[a];
            if (s >= S) selft = -1;
            const JSide J = {jside_s[0], jside_s[N3 ? 1 : 0], jside_s[N3 ? 2 : 0], jmin};
            for (int g0 = 0; g0 < trips; g0 += 64)
            {
                const int ng = min(64, trips - g0);
                unsigned long long word;
                if (replay)
                    word = g0 == 0 ? wfirst : (wi < hw ? D.hbits[((long)c * HBITS_WORDS + wi) * NT + lane] : 0ull);
                else
                {
                    /* bits are shifted in from the last trip down, so trip t lands on bit t */
                    unsigned lo = 0, hi = 0;
                    const int n0 = min(ng, 32);
                    int m = s + (g0 + ng - 1) * S;
                    if constexpr (NSCAN)
                    {
                        const float mx = live ? -2.0f * xi : 0.0f, my = live ? -2.0f * yi : 0.0f;
                        const float mz = live ? -2.0f * zi : 0.0f;
                        const float nim = live ? (xi * xi + yi * yi + zi * zi) - rscan2 : __builtin_inff();
                        const float4 *q = T.at(m);
#pragma unroll 4
                        for (int t = ng - 1; t >= 32; t--, q -= S)
                        {
                            const float4 v = *q;
                            const float d = __builtin_fmaf(v.z, mz, __builtin_fmaf(v.y, my, __builtin_fmaf(v.x, mx, v.w + nim)));
                            hi = __builtin_amdgcn_alignbit(hi, __builtin_bit_cast(unsigned, d), 31);
                        }
#pragma unroll 4
                        for (int t = n0 - 1; t >= 0; t--, q -= S)
                        {
                            const float4 v = *q;
                            const float d = __builtin_fmaf(v.z, mz, __builtin_fmaf(v.y, my, __builtin_fmaf(v.x, mx, v.w + nim)));
                            lo = __builtin_amdgcn_alignbit(lo, __builtin_bit_cast(unsigned, d), 31);
                        }
                    }
                    else
                    {
                        const R sx = live ? xi : pinf, sy = live ? yi : pinf, sz = live ? zi : pinf;
                        const auto *q = T.at(m);
#pragma unroll 4
                        for (int t = ng - 1; t >= 32; t--, q -= S)
                        {
                            const R dx = sx - q->x, dy = sy - q->y, dz = sz - q->z;
                            const R d = fma(dz, dz, fma(dy, dy, fma(dx, dx, -rscan2)));
                            hi = __builtin_amdgcn_alignbit(hi, sign_word(d), 31);
                        }
#pragma unroll 4
                        for (int t = n0 - 1; t >= 0; t--, q -= S)
                        {
                            const R dx = sx - q->x, dy = sy - q->y, dz = sz - q->z;
                            const R d = fma(dz, dz, fma(dy, dy, fma(dx, dx, -rscan2)));
                            lo = __builtin_amdgcn_alignbit(lo, sign_word(d), 31);
                        }
                    }
                    /* A lane without a live home atom keeps no hits: the lanes beyond the
                       stripes (s >= S) scan entries past the padded tile, whose stale LDS
                       contents can read as hits, and nothing of theirs may reach the
                       flushes, whose exact FP64 pairs turn a tile entry's list place into a
                       global slot index. */
                    word = live ? ((unsigned long long)hi << 32) | lo : 0ull;
                    /* the home atom meets itself on one trip of one stripe: clear that bit */
                    const int ds = selft - s - g0 * S;
                    if (ds >= 0 && ds < ng * S && ds % S == 0) word &= ~(1ull << (ds / S));
                    if (record && wi < HBITS_WORDS) D.hbits[((long)c * HBITS_WORDS + wi) * NT + lane] = word;
                }
                unsigned long long xw = 0ull;   /* FP32: trips of this word whose pairs are in the exact shell */
                if constexpr (CAT == CAT_EAM1)
                    flush_eam1<R, PASS, EN, LT>(D, T, word, s, g0, S, xi, yi, zi, Fpi, p1.rc2, xw, acc, lt);
                else if constexpr ((CAT == CAT_LJ || CAT == CAT_MORSE) && PASS == 0)
                {
                    /* the rounding band is met so seldom that the word's bits are walked again to
                       find its pairs rather than tracked pair by pair */
                    const unsigned long long w0 = word;
                    bool ex = false;
                    flush_pair2<R, CAT, EN, N3>(D, T, word, s, g0, S, xi, yi, zi, p1, ex, acc, J);
                    if constexpr (sizeof(R) == sizeof(float))
                        if (__ballot(ex)) xw = band_bits(T, w0, s, g0, S, xi, yi, zi, (R)(p1.rc2 * (1.0 - RX_BAND)));
                }
                else
                    while (__ballot(word != 0ull))
                    {
                        if (word)
                        {
                            const int t = __builtin_ctzll(word);
                            word &= word - 1ull;
                            const int m = tpos(s, g0 + t, S);
                            R x, y, z, fj;
                            T.load(m, x, y, z, fj);
                            const PairParam &pq = CAT == CAT_GEN ? ppi[T.k[m]] : p1;
                            if constexpr (sizeof(R) == sizeof(float))
                            {
                                /* the exact shell: the rounding band, and for EAM the table's steep run */
                                const R ex = xi - x, ey = yi - y, ez = zi - z;
                                const R r2 = ex * ex + ey * ey + ez * ez;
                                const int pc = CAT == CAT_GEN ? pq.cat : (CAT == CAT_EAM1 ? FMDX_POT_EAM : pq.cat);
                                const double rx = pc == FMDX_POT_EAM ? D.eam.rx2 : pq.rc2 * (1.0 - RX_BAND);
                                if (pc != FMDX_POT_NONE && r2 >= (R)rx)
                                {
                                    xw |= 1ull << t;
                                    continue;
                                }
                            }
                            pair_hit<R, PASS, CAT, Acc, EN>(D, pq, Fpi, fj, xi - x, yi - y, zi - z, acc);
                        }
                    }
                if constexpr (sizeof(R) == sizeof(float))
                {
                    if (__ballot(xw != 0ull)) flush_exact<PASS, CAT, EN, N3>(D, T, xw, s, g0, S, X, ppi, p1, acc, J);
                }
                wi++;
            }
            if constexpr (N3)
            {
                /* the round's forward candidates hand their partner sums over, each to the
                   entry of its slot for the direction it lies in; a second block of home atoms
                   (cells of more than 64) adds to what the first wrote */
                csync<LT>();
                if (h0 > 0) __threadfence();
                for (int m = jmin + lane; m < nt; m += NT)
                {
                    const int g = T.g[m];
                    const int lo = locate(g);
                    float *const pj = D.np + (long)(lo - 1) * nslot + nb_base[lo] + (g - nb_pre[lo]);
                    const float vx = jside_s[0][m], vy = jside_s[1][m], vz = jside_s[2][m];
                    if (h0 == 0)
                    {
                        pj[0] = vx;
                        pj[13 * nslot] = vy;
                        pj[26 * nslot] = vz;
                    }
                    else
                    {
                        pj[0] += vx;
                        pj[13 * nslot] += vy;
                        pj[26 * nslot] += vz;
                    }
                }
            }
        }
        if (record && lane == 0)
        {
            D.htrips[c] = wi <= HBITS_WORDS ? wi : -1;
            D.htile[c] = tiles;
        }
        csync<LT>();   /* the tile is reused for the stripe sums */
        red[0][lane] = acc[0].get();
        red[1][lane] = acc[1].get();
        red[2][lane] = acc[2].get();
        red[3][lane] = acc[3].get();
        csync<LT>();
        double ecell = 0.0;
        if (lane < nhc)
        {
            double sfx = 0, sfy = 0, sfz = 0, se = 0;
            for (int q = 0; q < S; q++)
            {
                const int l = lane + q * nhc;
                sfx += red[0][l]; sfy += red[1][l]; sfz += red[2][l]; se += red[3][l];
            }
            if (PASS == 1)
            {
                double Fp = 0.0;
                const int el = D.elem[ki];
                if (act && el >= 0)
                {
                    const EamDev &E = D.eam;
                    const double t = se * E.inv_drho;
                    int k = (int)t;
                    if (k >= E.nrho - 1 || k < 0 || !(se == se))
                    {
                        set_err(D, FMDX_DEV_ERR_RHO_RANGE);
                        k = k < 0 ? 0 : E.nrho - 2;
                    }
                    const double b = t - k;
                    const double4 q = E.Fc[(long)el * (E.nrho - 1) + k];
                    Fp = sp_der<double, double4>(q, b, E.inv_drho);
                    ecell = sp_val<double, double4>(q, b);
                }
                D.aux[is] = Fp;
                if constexpr (sizeof(R) == sizeof(float)) D.xq[is].w = (float)Fp;   /* for the force pass's staging */
            }
            else
            {
                const TtmDev &TT = D.ttm;
                if (act && TT.on)
                {
                    /* add_ttm_term_to_forces, forces.c:292-330 (z only for a 1 x 1 x n turi) */
                    /* the FP64 position, as k_ttm_index uses for the cell sums of the same atoms */
                    const int iz = (int)(D.z[is] / TT.tcellh_z) + 1;
                    if (iz >= 0 && iz < TT.ncz + 2)
                    {
                        if (TT.ncx * TT.ncy > 1)
                        {
                            int ix = (int)(D.x[is] / TT.tcellh_x), iy = (int)(D.y[is] / TT.tcellh_y);
                            ix = ix < 0 ? 0 : (ix >= TT.ncx ? TT.ncx - 1 : ix);
                            iy = iy < 0 ? 0 : (iy >= TT.ncy ? TT.ncy - 1 : iy);
                            const int a = (ix * TT.ncy + iy) * (TT.ncz + 2) + iz;
                            const double ximass = TT.xi[a] * D.mass[ki];
                            sfx += ximass * (D.vx[is] - TT.vcm[3 * a]);
                            sfy += ximass * (D.vy[is] - TT.vcm[3 * a + 1]);
                            sfz += ximass * (D.vz[is] - TT.vcm[3 * a + 2]);
                        }
                        else
                            sfz += TT.xi[iz] * D.mass[ki] * (D.vz[is] - TT.vcm[3 * iz + 2]);
                    }
                }
                if (act && !(__builtin_isfinite(sfx) && __builtin_isfinite(sfy) && __builtin_isfinite(sfz)))
                {
                    set_err(D, FMDX_DEV_ERR_FORCE);
                    note_first(&D.counters[4], (unsigned)c + 1u);
                }
                D.fx[is] = act ? sfx : 0.0;
                D.fy[is] = act ? sfy : 0.0;
                D.fz[is] = act ? sfz : 0.0;
                ecell = se;
            }
        }
        csync<LT>();
        red[0][lane] = ecell;
        csync<LT>();
        if (lane == 0)
            for (int l = 0; l < nhc; l++) cell_acc += red[0][l];
    }
    if (lane == 0)
    {
        if (PASS == 1) D.cell_emb[c] = cell_acc;
        else if (EN) D.cell_pe[c] = cell_acc;
    }
}

/* Cell visiting order of the single-system pair kernels.  Workgroups are
   dispatched to the 8 XCDs round-robin, so block b runs on XCD b % 8; each
   XCD gets a contiguous run of a sequence that walks the cells in y-strips
   of STRIP_W cells: strip by strip, plane x by plane x, then y and z.  A
   cell's atoms are staged by the 27 cells around it; in this order they are
   reused within about four strip-planes (STRIP_W * nc2 cells each, about
   2 MB for 4M-atom copper), which stay in the XCD's 4 MB L2, where the
   storage order (z fastest, then y, then x) would need three whole planes. */
#define STRIP_W 16
static __device__ __forceinline__ int xcd_cell(const KDesc &D, int b, int n)
{
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, i = b >> 3;
    const int k = x * q + min(x, r) + i;   /* position in the strip sequence */
    const int full = D.nc1 / STRIP_W;
    const int per_strip = D.nc0 * STRIP_W * D.nc2;
    int s, j, w;
    if (k < full * per_strip)
    {
        s = k / per_strip;
        j = k - s * per_strip;
        w = STRIP_W;
    }
    else
    {
        s = full;
        j = k - full * per_strip;
        w = D.nc1 - full * STRIP_W;
    }
    const int px = j / (w * D.nc2);
    const int rem = j - px * w * D.nc2;
    const int py = s * STRIP_W + rem / D.nc2;
    const int pz = rem % D.nc2;
    return (px * D.nc1 + py) * D.nc2 + pz;
}

template <typename R, int PASS, int CAT, bool EN, bool N3 = false>
__global__ __launch_bounds__(NT) void k_cell(KDesc D)
{
    b_cell<R, PASS, CAT, EN, false, N3>(D, xcd_cell(D, blockIdx.x, gridDim.x));
}

#if FMDX_EAM_LDS
/* FP32 single-element EAM pass with its spline table resident in LDS: one
   workgroup of LW waves per CU (the table, ~80 KB for the reference's 5000-
   point grids, leaves room for one), loading the table once; each wave then
   takes the cells of its XCD's run of the visiting order in turn, with a
   fixed stride (cells of one system cost about the same).  The density pass
   holds the density coefficients (c0..c3) per interval, the force pass the
   derivative knots of rho and phi. */
template <int PASS, bool EN>
__global__ __launch_bounds__(WG * LW) void k_cell_lds(KDesc D)
{
    extern __shared__ PairParam pp[];
    float4 *const tab = reinterpret_cast<float4 *>(reinterpret_cast<char *>(pp) + lt_offset(D.nkinds));
    const int nkk = D.nkinds * D.nkinds;
    for (int i = threadIdx.x; i < nkk; i += WG * LW) pp[i] = D.pairs[i];
    const float4 *const src = PASS == 1 ? D.eam.rhoc32 : D.eam.kn32;
    const int ntab = PASS == 1 ? D.eam.nr2 - 1 : D.eam.nr2;
    for (int i = threadIdx.x; i < ntab; i += WG * LW) tab[i] = src[i];
    __syncthreads();
    const int x = blockIdx.x & 7;                 /* the XCD this workgroup runs on */
    const int q = (int)(gridDim.x >> 3) * LW;     /* waves per XCD */
    for (int i = (int)(blockIdx.x >> 3) * LW + (int)(threadIdx.x / WG);; i += q)
    {
        const int v = x + 8 * i;                  /* a block index of the one-cell-per-workgroup launch */
        if (v >= D.ncell) break;
        b_cell<float, PASS, CAT_EAM1, EN, true>(D, xcd_cell(D, v, D.ncell));
    }
}
#endif

/* LOOP: the cells of a compacted list beyond the grid of the one-cell-per-workgroup
   launch (M.skip), when the list has outgrown it since the host last sized it */
template <typename R, int PASS, int CAT, bool EN, bool N3 = false, bool LOOP = false>
__global__ __launch_bounds__(NT) void kb_cell(BatchMap M)
{
    each_item<LOOP>(M, [&](const KDesc &D, int c) { b_cell<R, PASS, CAT, EN, false, N3>(D, c); });
}

/* N3: a slot's force is its own home-side sum plus the partner sums left for it by the
   home cells of its 13 backward directions, in direction order (FP64 sums of FP32
   terms); a direction without a home cell (an open face) or with an empty one adds
   nothing (an empty home cell writes no partner sums) */
static __device__ void b_fold(const KDesc &D, int blk)
{
    if (!D.np) return;
    const long nslot = (long)D.ncell * D.cap;
    each_slot(D, blk, [&](long slot, int c) {
        const int cz = c % D.nc2, t = c / D.nc2, cy = t % D.nc1, cx = t / D.nc1;
        double fx = D.fx[slot], fy = D.fy[slot], fz = D.fz[slot];
        for (int f = 0; f < 13; f++)
        {
            int ox, oy, oz;
            n3_offset(f, ox, oy, oz);
            int hx = cx - ox, hy = cy - oy, hz = cz - oz;
            bool ok = true;
            if (hx < 0) { if (D.pbc0) hx += D.nc0; else ok = false; }
            if (hy < 0) { if (D.pbc1) hy += D.nc1; else ok = false; }
            else if (hy >= D.nc1) { if (D.pbc1) hy -= D.nc1; else ok = false; }
            if (hz < 0) { if (D.pbc2) hz += D.nc2; else ok = false; }
            else if (hz >= D.nc2) { if (D.pbc2) hz -= D.nc2; else ok = false; }
            if (!ok || D.cnt[((long)hx * D.nc1 + hy) * D.nc2 + hz] == 0) continue;
            const float *const pj = D.np + (long)f * nslot + slot;
            fx += (double)pj[0];
            fy += (double)pj[13 * nslot];
            fz += (double)pj[26 * nslot];
        }
        if (!(__builtin_isfinite(fx) && __builtin_isfinite(fy) && __builtin_isfinite(fz)))
        {
            set_err(D, FMDX_DEV_ERR_FORCE);
            note_first(&D.counters[4], (unsigned)c + 1u);
        }
        D.fx[slot] = fx;
        D.fy[slot] = fy;
        D.fz[slot] = fz;
    });
}
__global__ __launch_bounds__(256) void k_fold(KDesc D) { b_fold(D, blockIdx.x); }
__global__ __launch_bounds__(256) void kb_fold(BatchMap M)
{
    each_item<true>(M, [&](const KDesc &D, int blk) { b_fold(D, blk); });
}
/* FP32 positions relative to the origin of the home cell, from the FP64 positions */
static __device__ void b_relpos(const KDesc &D, int blk)
{
    each_slot(D, blk, [&](long slot, int c) {
        const int cz = c % D.nc2, t = c / D.nc2, cy = t % D.nc1, cx = t / D.nc1;
        D.xq[slot] = rel_pos(D, cx, cy, cz, D.x[slot], D.y[slot], D.z[slot]);
    });
}
__global__ __launch_bounds__(256) void k_relpos(KDesc D) { b_relpos(D, blockIdx.x); }
__global__ __launch_bounds__(256) void kb_relpos(BatchMap M)
{
    each_item<true>(M, [&](const KDesc &D, int blk) { b_relpos(D, blk); });
}

static inline size_t pp_bytes(const KDesc &D) { return sizeof(PairParam) * D.nkinds * D.nkinds; }

int force_threads_per_cell() { return NT; }

int force_category(const KDesc &D)
{
    if (D.single_cat == FMDX_POT_EAM && D.eam.ne == 1 && D.eam.rf) return CAT_EAM1;
    if (D.nkinds == 1 && D.single_cat == FMDX_POT_LJ) return CAT_LJ;
    if (D.nkinds == 1 && D.single_cat == FMDX_POT_MORSE) return CAT_MORSE;
    return CAT_GEN;
}

#if FMDX_EAM_LDS
/* LDS of the LDS-table kernels for system D: static (LW tiles) plus dynamic (pair
   parameters and the table); 0 when the kernels cannot hold D's table.  Both passes
   take the same decision, so that the force pass replays the density pass's tiles. */
static size_t lds_table_bytes(const KDesc &D)
{
    if (!D.eam.kn32 || !D.eam.rhoc32) return 0;
    static int limit = -1, ncu = 0;
    static size_t stat = 0;
    if (limit < 0)
    {
        int dev = 0;
        hipFuncAttributes a1, a0, a0e;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipFuncGetAttributes(&a1, reinterpret_cast<const void *>(&k_cell_lds<1, true>)) != hipSuccess ||
            hipFuncGetAttributes(&a0, reinterpret_cast<const void *>(&k_cell_lds<0, false>)) != hipSuccess ||
            hipFuncGetAttributes(&a0e, reinterpret_cast<const void *>(&k_cell_lds<0, true>)) != hipSuccess)
        {
            limit = 0;
            return 0;
        }
        stat = a1.sharedSizeBytes > a0.sharedSizeBytes ? a1.sharedSizeBytes : a0.sharedSizeBytes;
        if (a0e.sharedSizeBytes > stat) stat = a0e.sharedSizeBytes;
    }
    const size_t dyn = lt_offset(D.nkinds) + (size_t)D.eam.nr2 * sizeof(float4);
    return ncu >= 8 && stat + dyn <= (size_t)limit ? dyn : 0;
}

static int lds_grid()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return ncu & ~7;   /* one workgroup per CU, a whole number per XCD */
}

/* dynamic LDS beyond the default limit has to be granted per kernel first */
template <int PASS, bool EN>
static void launch_lds(const KDesc &D, size_t dyn, hipStream_t st)
{
    static size_t granted = 0;
    if (dyn > granted)
    {
        /* on failure the error stays for the caller's hipGetLastError */
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_cell_lds<PASS, EN>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
            return;
        granted = dyn;
    }
    hipLaunchKernelGGL((k_cell_lds<PASS, EN>), dim3(lds_grid()), dim3(WG * LW), dyn, st, D);
}
#endif

template <typename R, int PASS, bool EN>
static void launch_cell_en(const KDesc &D, hipStream_t st)
{
    const dim3 g(D.ncell), b(NT);
    const size_t lds = pp_bytes(D);
    switch (force_category(D))
    {
        case CAT_EAM1:
#if FMDX_EAM_LDS
            if constexpr (sizeof(R) == sizeof(float))
                if (const size_t dyn = lds_table_bytes(D))
                {
                    launch_lds<PASS, EN>(D, dyn, st);
                    break;
                }
#endif
            hipLaunchKernelGGL((k_cell<R, PASS, CAT_EAM1, EN>), g, b, lds, st, D);
            break;
        case CAT_LJ:
        case CAT_MORSE:
            if constexpr (sizeof(R) == sizeof(float) && PASS == 0 && FMDX_N3 && FMDX_WAVES_PER_CELL == 1)
                if (D.np)
                {
                    if (force_category(D) == CAT_LJ) hipLaunchKernelGGL((k_cell<R, PASS, CAT_LJ, EN, true>), g, b, lds, st, D);
                    else hipLaunchKernelGGL((k_cell<R, PASS, CAT_MORSE, EN, true>), g, b, lds, st, D);
                    hipLaunchKernelGGL(k_fold, dim3((D.ncell + CPB - 1) / CPB), dim3(TPB), 0, st, D);
                    break;
                }
            if (force_category(D) == CAT_LJ) hipLaunchKernelGGL((k_cell<R, PASS, CAT_LJ, EN>), g, b, lds, st, D);
            else hipLaunchKernelGGL((k_cell<R, PASS, CAT_MORSE, EN>), g, b, lds, st, D);
            break;
        default: hipLaunchKernelGGL((k_cell<R, PASS, CAT_GEN, EN>), g, b, lds, st, D); break;
    }
}

/* each pair evaluated once for this system: mixed precision, single-kind LJ or Morse */
bool n3_wanted(const KDesc &D, int precision)
{
    const int cat = force_category(D);
    return FMDX_N3 && FMDX_WAVES_PER_CELL == 1 && precision == FMDX_PREC_MIXED && (cat == CAT_LJ || cat == CAT_MORSE);
}

/* the density pass has one form; the force pass takes the pair energy when D.energy is set */
template <typename R, int PASS>
static void launch_cell(const KDesc &D, hipStream_t st)
{
    if (PASS == 1 || D.energy) launch_cell_en<R, PASS, true>(D, st);
    else launch_cell_en<R, PASS, false>(D, st);
}

/* pass 1 (EAM density + F'(rho)) and pass 0 (forces + energy), launched separately so that
   callers can time each kernel with events on the stream */
hipError_t launch_pass1(const KDesc &D, hipStream_t st, int precision, bool relpos)
{
    if (precision == FMDX_PREC_MIXED)
    {
        if (relpos) hipLaunchKernelGGL(k_relpos, dim3((D.ncell + CPB - 1) / CPB), dim3(TPB), 0, st, D);
        if (D.has_eam) launch_cell<float, 1>(D, st);
        else (void)launch_fill(D.cell_emb, D.ncell, 0.0, st);
    }
    else
    {
        if (D.has_eam) launch_cell<double, 1>(D, st);
        else (void)launch_fill(D.cell_emb, D.ncell, 0.0, st);
    }
    return hipGetLastError();
}

hipError_t launch_pass0(const KDesc &D, hipStream_t st, int precision)
{
    if (precision == FMDX_PREC_MIXED) launch_cell<float, 0>(D, st);
    else launch_cell<double, 0>(D, st);
    return hipGetLastError();
}

hipError_t launch_forces(const KDesc &D, hipStream_t st, int precision, bool relpos)
{
    hipError_t e = launch_pass1(D, st, precision, relpos);
    if (e != hipSuccess) return e;
    e = launch_pass0(D, st, precision);
    if (e != hipSuccess) return e;
    return D.energy ? launch_reduce_pe(D, st) : hipSuccess;
}

/* one category's pass over a list: one workgroup per cell, and, on a compacted list whose
   grid is shorter than the full list, a small looping launch for the cells beyond that grid
   (it exits at once when there are none) */
#define GRID_MORE 64
template <typename R, int PASS, int CAT, bool EN, bool N3>
static void launchb_cat(const BatchMap &mc, dim3 g, int full, size_t ppb, hipStream_t st)
{
    hipLaunchKernelGGL((kb_cell<R, PASS, CAT, EN, N3>), g, dim3(NT), ppb, st, mc);
    if (mc.count && (int)g.x < full)
    {
        BatchMap mo = mc;
        mo.skip = (int)g.x;
        hipLaunchKernelGGL((kb_cell<R, PASS, CAT, EN, N3, true>), dim3(GRID_MORE), dim3(NT), ppb, st, mo);
    }
}

/* the pass over every cell of the members (lists, n), or, compacted, over the occupied
   ones (occ, occ_n[at + cat]) with the host-sized grids */
template <typename R, int PASS, bool EN>
static void launchb_cells(const BatchPlan &P, hipStream_t st, const int2 *const *lists, const int *n,
                          int2 *const *occ, const int *grid, int at)
{
    const size_t ppb = sizeof(PairParam) * P.max_nkinds * P.max_nkinds;
    constexpr bool N3OK = sizeof(R) == sizeof(float) && PASS == 0 && FMDX_N3 && FMDX_WAVES_PER_CELL == 1;
    for (int cat = 0; cat < 4; cat++)
    {
        if (n[cat] == 0) continue;
        BatchMap mc = {P.descs, lists[cat], nullptr, 0};
        dim3 g(n[cat]);
        if (P.compact)
        {
            mc = {P.descs, occ[cat], P.occ_n + at + cat, 0};
            g = dim3((unsigned)grid[cat]);
        }
        switch (cat)
        {
            case CAT_EAM1: launchb_cat<R, PASS, CAT_EAM1, EN, false>(mc, g, n[cat], ppb, st); break;
            case CAT_LJ:
                if (N3OK && P.n3) launchb_cat<R, PASS, CAT_LJ, EN, N3OK>(mc, g, n[cat], ppb, st);
                else launchb_cat<R, PASS, CAT_LJ, EN, false>(mc, g, n[cat], ppb, st);
                break;
            case CAT_MORSE:
                if (N3OK && P.n3) launchb_cat<R, PASS, CAT_MORSE, EN, N3OK>(mc, g, n[cat], ppb, st);
                else launchb_cat<R, PASS, CAT_MORSE, EN, false>(mc, g, n[cat], ppb, st);
                break;
            default: launchb_cat<R, PASS, CAT_GEN, EN, false>(mc, g, n[cat], ppb, st); break;
        }
    }
    /* each pair once: the partner sums folded into the forces (members without them pass) */
    if constexpr (sizeof(R) == sizeof(float) && PASS == 0 && FMDX_N3 && FMDX_WAVES_PER_CELL == 1)
        if (P.n3)
        {
            BatchMap ms = {P.descs, P.slot_items, nullptr, 0};
            dim3 gs(P.n_slot);
            if (P.compact)
            {
                ms = {P.descs, P.occ_slot, P.occ_n, 0};
                gs = dim3((unsigned)P.g_slot);
            }
            hipLaunchKernelGGL(kb_fold, gs, dim3(TPB), 0, st, ms);
        }
}

/* every system of a batch: relative positions (mixed), the density pass over
   the EAM systems, the force pass and, with `energy`, the pair energies and
   their sums; one launch per pair category */
template <bool EN>
static void launchb_passes(const BatchPlan &P, hipStream_t st, int precision, bool relpos)
{
    if (precision == FMDX_PREC_MIXED)
    {
        BatchMap ms = {P.descs, P.slot_items, nullptr, 0};
        if (relpos) hipLaunchKernelGGL(kb_relpos, dim3(P.n_slot), dim3(TPB), 0, st, ms);
        launchb_cells<float, 1, true>(P, st, P.cell1, P.n_cell1, P.occ_cell1, P.g_cell1, 5);
        launchb_cells<float, 0, EN>(P, st, P.cell0, P.n_cell0, P.occ_cell0, P.g_cell0, 1);
    }
    else
    {
        launchb_cells<double, 1, true>(P, st, P.cell1, P.n_cell1, P.occ_cell1, P.g_cell1, 5);
        launchb_cells<double, 0, EN>(P, st, P.cell0, P.n_cell0, P.occ_cell0, P.g_cell0, 1);
    }
}

hipError_t launchb_forces(const BatchPlan &P, hipStream_t st, int precision, int energy, bool relpos)
{
    if (P.nsys == 0) return hipSuccess;
    if (energy) launchb_passes<true>(P, st, precision, relpos);
    else launchb_passes<false>(P, st, precision, relpos);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return energy ? launchb_reduce_pe(P, st) : hipSuccess;
}
