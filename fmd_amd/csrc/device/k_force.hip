/*
 * k_force.hip -- 27-neighbour-cell pair kernels for LJ, Morse and two-pass EAM.
 *
 * One 64-lane wavefront owns one home cell.  Its lanes are split into
 * (home atom a, stripe s): with n home atoms each atom gets S = 64/n lanes,
 * and lane (a,s) tests every S-th atom of the 27 neighbour cells.  The scan
 * is wave-uniform (same trip count on every lane); a lane that finds a
 * partner inside the cutoff only records it in an LDS hit buffer, and the
 * spline / exp work runs in a separate flush loop over those buffers so that
 * it is never executed on behalf of a lane without a hit.  Partial results
 * of the S stripes are combined in a fixed order, so every atom's sum is
 * deterministic.
 *
 * Full shell: each atom accumulates only its own force / density, which
 * removes the reference's Newton-3 writes into neighbour cells (the OpenMP
 * race of lj.c:37-87, eam.c:38-142) and the ghost fold-back
 * (md-ghost.c:788-825).  Periodic images are formed by shifting the home
 * position by the box length of the wrapped neighbour cell.
 *
 * Pair set and formulas:
 *   LJ      lj.h:26-58        F += 12*(4eps)/r^2 * ((s/r)^12 - (s/r)^6/2) r,   U = 4eps((s/r)^12-(s/r)^6)
 *   Morse   morse.h:26-58     F += 2 a D0/r (e^2 - e) r, e = exp(-a(r-r0)),    U = D0(e^2 - 2e)
 *   EAM p1  eam.h:270-379     rho_i = sum_j rho_elem(j)(r^2)   then F'(rho), F(rho)  eam.c:160-223
 *   EAM p0  eam.h:29-228      F_i -= 2(F'_i rho'_elem(j) + F'_j rho'_elem(i) + phi'_ij) r
 * with every tabulated function a natural cubic spline on the reference's
 * uniform r^2 (or rho) grid, stored as per-interval polynomial coefficients.
 * The pair energy of a pair is accumulated half on each side.
 */
#include "fmdx_internal.h"

#define WG 64
#define HB 32

static __device__ __forceinline__ unsigned long long wave_any(bool p)
{
    return __ballot(p);
}

/* spline value and derivative (w.r.t. x) on interval k with b = x/h - k */
static __device__ __forceinline__ double sp_val(const double4 &q, double b)
{
    return q.x + b * (q.y + b * (q.z + b * q.w));
}
static __device__ __forceinline__ double sp_der(const double4 &q, double b, double inv_h)
{
    return (q.y + b * (2.0 * q.z + 3.0 * b * q.w)) * inv_h;
}
static __device__ __forceinline__ float sp_val32(const float4 &q, float b)
{
    return q.x + b * (q.y + b * (q.z + b * q.w));
}
static __device__ __forceinline__ float sp_der32(const float4 &q, float b, float inv_h)
{
    return (q.y + b * (2.0f * q.z + 3.0f * b * q.w)) * inv_h;
}

struct Acc
{
    double fx, fy, fz, e;
};

/* per-wave LDS scratch: the hit buffer during the scan, the stripe sums afterwards */
union ShScratch
{
    uint16_t hb[HB][WG];
    double red[4][WG];
};

/* Kahan-compensated FP32 accumulator */
struct KahanF
{
    float s, c;
    __device__ void add(float v)
    {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

/* ---------------------------------------------------------------- FP64 */

template <int PASS>
static __device__ __forceinline__ void hit64(const KDesc &D, const PairParam *pp, int nk, int ki, double Fpi,
                                             double dx, double dy, double dz, long js, Acc &acc)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    const int kj = D.kind[js];
    const PairParam &p = pp[ki * nk + kj];
    if (!(r2 < p.rc2)) return;
    if (PASS == 1)
    {
        if (p.cat != FMDX_POT_EAM) return;
        const EamDev &E = D.eam;
        double t = r2 * E.inv_dr2;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        double b = t - k;
        acc.e += sp_val(E.rhoc[(long)p.jloc * (E.nr2 - 1) + k], b);
        return;
    }
    if (p.cat == FMDX_POT_EAM)
    {
        const EamDev &E = D.eam;
        double t = r2 * E.inv_dr2;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        double b = t - k;
        const long m = E.nr2 - 1;
        double rho_ip = sp_der(E.rhoc[(long)p.iloc * m + k], b, E.inv_dr2);
        double rho_jp = (p.jloc == p.iloc) ? rho_ip : sp_der(E.rhoc[(long)p.jloc * m + k], b, E.inv_dr2);
        const double4 q = E.phic[(long)(p.iloc * E.ne + p.jloc) * m + k];
        double phid = sp_der(q, b, E.inv_dr2);
        double Fpj = D.aux[js];
        double mag = 2.0 * (Fpi * rho_jp + Fpj * rho_ip + phid);
        acc.fx -= mag * dx;
        acc.fy -= mag * dy;
        acc.fz -= mag * dz;
        acc.e += 0.5 * sp_val(q, b);
    }
    else if (p.cat == FMDX_POT_LJ)
    {
        double inv_r2 = 1.0 / r2;
        double inv_rs2 = p.p1 * inv_r2;
        double inv_rs6 = inv_rs2 * inv_rs2 * inv_rs2;
        double inv_rs12 = inv_rs6 * inv_rs6;
        double factor = 12.0 * p.p0 * inv_r2 * (inv_rs12 - 0.5 * inv_rs6);
        acc.fx += factor * dx;
        acc.fy += factor * dy;
        acc.fz += factor * dz;
        acc.e += 0.5 * p.p0 * (inv_rs12 - inv_rs6);
    }
    else if (p.cat == FMDX_POT_MORSE)
    {
        double r = sqrt(r2);
        double e1 = exp(-p.p1 * (r - p.p2));
        double e2 = e1 * e1;
        double factor = 2.0 * p.p1 * p.p0 / r * (e2 - e1);
        acc.fx += factor * dx;
        acc.fy += factor * dy;
        acc.fz += factor * dz;
        acc.e += 0.5 * p.p0 * (e2 - 2.0 * e1);
    }
}

template <int PASS>
static __device__ void b_pairs64(const KDesc &D, const int c)
{
    const int ag = D.ag;
    const TtmDev &T = D.ttm;
    if (PASS == 1 && !D.has_eam)
    {
        if (threadIdx.x == 0) D.cell_emb[c] = 0.0;
        return;
    }
    __shared__ int nb_base[27], nb_cnt[27];
    __shared__ double nb_sx[27], nb_sy[27], nb_sz[27];
    __shared__ ShScratch sm;
    extern __shared__ PairParam pp[];   /* nkinds^2 entries, sized at launch */
    uint16_t(&hb)[HB][WG] = sm.hb;
    double(&red)[4][WG] = sm.red;

    const int lane = threadIdx.x;
    const int nkk = D.nkinds * D.nkinds;
    for (int i = lane; i < nkk; i += WG) pp[i] = D.pairs[i];
    if (lane < 27)
    {
        int cz = c % D.nc2, t = c / D.nc2, cy = t % D.nc1, cx = t / D.nc1;
        int jx = cx + lane / 9 - 1, jy = cy + (lane / 3) % 3 - 1, jz = cz + lane % 3 - 1;
        double sx = 0, sy = 0, sz = 0;
        bool ok = true;
        if (jx < 0) { if (D.pbc0) { jx += D.nc0; sx = -D.L0; } else ok = false; }
        else if (jx >= D.nc0) { if (D.pbc0) { jx -= D.nc0; sx = D.L0; } else ok = false; }
        if (jy < 0) { if (D.pbc1) { jy += D.nc1; sy = -D.L1; } else ok = false; }
        else if (jy >= D.nc1) { if (D.pbc1) { jy -= D.nc1; sy = D.L1; } else ok = false; }
        if (jz < 0) { if (D.pbc2) { jz += D.nc2; sz = -D.L2; } else ok = false; }
        else if (jz >= D.nc2) { if (D.pbc2) { jz -= D.nc2; sz = D.L2; } else ok = false; }
        int jc = ok ? (jx * D.nc1 + jy) * D.nc2 + jz : 0;
        nb_base[lane] = jc * D.cap;
        nb_cnt[lane] = ok ? D.cnt[jc] : 0;
        nb_sx[lane] = sx; nb_sy[lane] = sy; nb_sz[lane] = sz;
    }
    __syncthreads();
    const int nh = D.cnt[c];
    double cell_acc = 0.0;
    const double rcut2 = PASS == 1 ? D.rceam2 : D.rcmax2;

    for (int h0 = 0; h0 < nh; h0 += WG)
    {
        const int nhc = min(WG, nh - h0);
        const int S = WG / nhc;
        const int a = lane % nhc, s = lane / nhc;
        const long is = (long)c * D.cap + h0 + a;
        double xi = 0, yi = 0, zi = 0, Fpi = 0;
        int ki = 0;
        bool act = false;
        if (s < S)
        {
            xi = D.x[is]; yi = D.y[is]; zi = D.z[is];
            ki = D.kind[is];
            act = (ag < 0 || D.grp[is] == ag);
            if (PASS == 1 && D.elem[ki] < 0) act = false;
            if (PASS == 0 && D.has_eam) Fpi = D.aux[is];
        }
        Acc acc = {0, 0, 0, 0};
        int fill = 0;
        for (int k = 0; k < 27; k++)
        {
            const int nk = nb_cnt[k];
            if (nk == 0) continue;
            const long base = nb_base[k];
            const double xs = xi - nb_sx[k], ys = yi - nb_sy[k], zs = zi - nb_sz[k];
            const int trips = (nk + S - 1) / S;
            for (int t = 0; t < trips; t++)
            {
                const int j = s + t * S;
                bool hit = false;
                if (act && j < nk && !(k == 13 && j == h0 + a))
                {
                    const long js = base + j;
                    double dx = xs - D.x[js], dy = ys - D.y[js], dz = zs - D.z[js];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    hit = r2 < rcut2;
                    if (hit && ag >= 0 && D.grp[js] != ag) hit = false;
                }
                if (hit) hb[fill++][lane] = (uint16_t)((k << 8) | j);
                if (wave_any(fill == HB))
                {
                    for (int h = 0; wave_any(h < fill); h++)
                    {
                        if (h < fill)
                        {
                            unsigned e = hb[h][lane];
                            int kk = e >> 8, jj = e & 255;
                            long js = nb_base[kk] + jj;
                            double dx = (xi - nb_sx[kk]) - D.x[js];
                            double dy = (yi - nb_sy[kk]) - D.y[js];
                            double dz = (zi - nb_sz[kk]) - D.z[js];
                            hit64<PASS>(D, pp, D.nkinds, ki, Fpi, dx, dy, dz, js, acc);
                        }
                    }
                    fill = 0;
                }
            }
        }
        for (int h = 0; wave_any(h < fill); h++)
        {
            if (h < fill)
            {
                unsigned e = hb[h][lane];
                int kk = e >> 8, jj = e & 255;
                long js = nb_base[kk] + jj;
                double dx = (xi - nb_sx[kk]) - D.x[js];
                double dy = (yi - nb_sy[kk]) - D.y[js];
                double dz = (zi - nb_sz[kk]) - D.z[js];
                hit64<PASS>(D, pp, D.nkinds, ki, Fpi, dx, dy, dz, js, acc);
            }
        }
        __syncthreads();   /* the hit buffer and the reduction scratch share LDS */
        red[0][lane] = acc.fx; red[1][lane] = acc.fy; red[2][lane] = acc.fz; red[3][lane] = acc.e;
        __syncthreads();
        double ecell = 0.0;
        if (lane < nhc)
        {
            double sfx = 0, sfy = 0, sfz = 0, se = 0;
            for (int q = 0; q < S; q++)
            {
                int l = lane + q * nhc;
                sfx += red[0][l]; sfy += red[1][l]; sfz += red[2][l]; se += red[3][l];
            }
            if (PASS == 1)
            {
                double Fp = 0.0;
                if (act)
                {
                    const EamDev &E = D.eam;
                    int el = D.elem[ki];
                    double t = se * E.inv_drho;
                    int k = (int)t;
                    if (k >= E.nrho - 1 || k < 0 || !(se == se))
                    {
                        set_err(D, FMDX_DEV_ERR_RHO_RANGE);
                        k = k < 0 ? 0 : E.nrho - 2;
                    }
                    double b = t - k;
                    const double4 q = E.Fc[(long)el * (E.nrho - 1) + k];
                    Fp = sp_der(q, b, E.inv_drho);
                    ecell = sp_val(q, b);
                }
                D.aux[is] = Fp;
            }
            else
            {
                if (act && T.on)
                {
                    int iz = (int)(zi / T.tcellh_z) + 1;
                    if (iz >= 0 && iz < T.ncz + 2)
                        sfz += T.xi[iz] * D.mass[ki] * (D.vz[is] - T.vcm[3 * iz + 2]);
                }
                D.fx[is] = act ? sfx : 0.0;
                D.fy[is] = act ? sfy : 0.0;
                D.fz[is] = act ? sfz : 0.0;
                ecell = se;
            }
        }
        __syncthreads();
        red[0][lane] = ecell;
        __syncthreads();
        if (lane == 0)
            for (int l = 0; l < nhc; l++) cell_acc += red[0][l];
        __syncthreads();
    }
    if (lane == 0)
    {
        if (PASS == 1) D.cell_emb[c] = cell_acc;
        else D.cell_pe[c] = cell_acc;
    }
}

template <int PASS>
__global__ __launch_bounds__(WG) void k_pairs64(KDesc D) { b_pairs64<PASS>(D, blockIdx.x); }

template <int PASS>
__global__ __launch_bounds__(WG) void kb_pairs64(BatchMap M)
{
    const int2 it = M.items[blockIdx.x];
    const KDesc D = M.descs[it.x];
    b_pairs64<PASS>(D, it.y);
}

/* ------------------------------------------------------------- mixed */

/* FP32 pair arithmetic on displacements taken from the FP64 positions,
   Kahan-compensated FP32 accumulation, FP64 results */
template <int PASS>
static __device__ __forceinline__ void hit32(const KDesc &D, const PairParam *pp, int nk, int ki, float Fpi,
                                             float dx, float dy, float dz, long js,
                                             KahanF &ax, KahanF &ay, KahanF &az, KahanF &ae)
{
    const float r2 = dx * dx + dy * dy + dz * dz;
    const int kj = D.kind[js];
    const PairParam &p = pp[ki * nk + kj];
    if (!(r2 < (float)p.rc2)) return;
    const EamDev &E = D.eam;
    if (PASS == 1)
    {
        if (p.cat != FMDX_POT_EAM) return;
        float t = r2 * (float)E.inv_dr2;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        float b = t - (float)k;
        ae.add(sp_val32(E.rhoc32[(long)p.jloc * (E.nr2 - 1) + k], b));
        return;
    }
    if (p.cat == FMDX_POT_EAM)
    {
        float inv = (float)E.inv_dr2;
        float t = r2 * inv;
        int k = (int)t;
        if (k > E.nr2 - 2) k = E.nr2 - 2;
        float b = t - (float)k;
        const long m = E.nr2 - 1;
        float rho_ip = sp_der32(E.rhoc32[(long)p.iloc * m + k], b, inv);
        float rho_jp = (p.jloc == p.iloc) ? rho_ip : sp_der32(E.rhoc32[(long)p.jloc * m + k], b, inv);
        const float4 q = E.phic32[(long)(p.iloc * E.ne + p.jloc) * m + k];
        float phid = sp_der32(q, b, inv);
        float Fpj = (float)D.aux[js];
        float mag = 2.0f * (Fpi * rho_jp + Fpj * rho_ip + phid);
        ax.add(-mag * dx);
        ay.add(-mag * dy);
        az.add(-mag * dz);
        ae.add(0.5f * sp_val32(q, b));
    }
    else if (p.cat == FMDX_POT_LJ)
    {
        float inv_r2 = 1.0f / r2;
        float inv_rs2 = (float)p.p1 * inv_r2;
        float inv_rs6 = inv_rs2 * inv_rs2 * inv_rs2;
        float inv_rs12 = inv_rs6 * inv_rs6;
        float factor = 12.0f * (float)p.p0 * inv_r2 * (inv_rs12 - 0.5f * inv_rs6);
        ax.add(factor * dx);
        ay.add(factor * dy);
        az.add(factor * dz);
        ae.add(0.5f * (float)p.p0 * (inv_rs12 - inv_rs6));
    }
    else if (p.cat == FMDX_POT_MORSE)
    {
        float r = sqrtf(r2);
        float e1 = expf(-(float)p.p1 * (r - (float)p.p2));
        float e2 = e1 * e1;
        float factor = 2.0f * (float)p.p1 * (float)p.p0 / r * (e2 - e1);
        ax.add(factor * dx);
        ay.add(factor * dy);
        az.add(factor * dz);
        ae.add(0.5f * (float)p.p0 * (e2 - 2.0f * e1));
    }
}

template <int PASS>
static __device__ void b_pairs32(const KDesc &D, const int c)
{
    const int ag = D.ag;
    const TtmDev &T = D.ttm;
    if (PASS == 1 && !D.has_eam)
    {
        if (threadIdx.x == 0) D.cell_emb[c] = 0.0;
        return;
    }
    __shared__ int nb_base[27], nb_cnt[27];
    __shared__ float nb_ox[27], nb_oy[27], nb_oz[27];   /* neighbour-cell origin minus home origin */
    __shared__ ShScratch sm;
    extern __shared__ PairParam pp[];   /* nkinds^2 entries, sized at launch */
    uint16_t(&hb)[HB][WG] = sm.hb;
    double(&red)[4][WG] = sm.red;

    const int lane = threadIdx.x;
    const int nkk = D.nkinds * D.nkinds;
    for (int i = lane; i < nkk; i += WG) pp[i] = D.pairs[i];
    const int cz = c % D.nc2, ct = c / D.nc2, cy = ct % D.nc1, cx = ct / D.nc1;
    if (lane < 27)
    {
        int dxn = lane / 9 - 1, dyn = (lane / 3) % 3 - 1, dzn = lane % 3 - 1;
        int jx = cx + dxn, jy = cy + dyn, jz = cz + dzn;
        double sx = 0, sy = 0, sz = 0;
        bool ok = true;
        if (jx < 0) { if (D.pbc0) { jx += D.nc0; sx = -D.L0; } else ok = false; }
        else if (jx >= D.nc0) { if (D.pbc0) { jx -= D.nc0; sx = D.L0; } else ok = false; }
        if (jy < 0) { if (D.pbc1) { jy += D.nc1; sy = -D.L1; } else ok = false; }
        else if (jy >= D.nc1) { if (D.pbc1) { jy -= D.nc1; sy = D.L1; } else ok = false; }
        if (jz < 0) { if (D.pbc2) { jz += D.nc2; sz = -D.L2; } else ok = false; }
        else if (jz >= D.nc2) { if (D.pbc2) { jz -= D.nc2; sz = D.L2; } else ok = false; }
        int jc = ok ? (jx * D.nc1 + jy) * D.nc2 + jz : 0;
        nb_base[lane] = jc * D.cap;
        nb_cnt[lane] = ok ? D.cnt[jc] : 0;
        /* origin of the image of cell jc relative to the home cell origin (FP64, then rounded) */
        nb_ox[lane] = (float)((jx * D.h0 + sx) - cx * D.h0);
        nb_oy[lane] = (float)((jy * D.h1 + sy) - cy * D.h1);
        nb_oz[lane] = (float)((jz * D.h2 + sz) - cz * D.h2);
    }
    __syncthreads();
    const int nh = D.cnt[c];
    double cell_acc = 0.0;
    const float rcut2 = (float)(PASS == 1 ? D.rceam2 : D.rcmax2);
    const double ox = cx * D.h0, oy = cy * D.h1, oz = cz * D.h2;

    for (int h0 = 0; h0 < nh; h0 += WG)
    {
        const int nhc = min(WG, nh - h0);
        const int S = WG / nhc;
        const int a = lane % nhc, s = lane / nhc;
        const long is = (long)c * D.cap + h0 + a;
        float xi = 0, yi = 0, zi = 0, Fpi = 0;
        int ki = 0;
        bool act = false;
        if (s < S)
        {
            xi = D.xr[is]; yi = D.yr[is]; zi = D.zr[is];
            ki = D.kind[is];
            act = (ag < 0 || D.grp[is] == ag);
            if (PASS == 1 && D.elem[ki] < 0) act = false;
            if (PASS == 0 && D.has_eam) Fpi = (float)D.aux[is];
        }
        KahanF ax = {0, 0}, ay = {0, 0}, az = {0, 0}, ae = {0, 0};
        int fill = 0;
        for (int k = 0; k < 27; k++)
        {
            const int nk = nb_cnt[k];
            if (nk == 0) continue;
            const long base = nb_base[k];
            const float xs = xi - nb_ox[k], ys = yi - nb_oy[k], zs = zi - nb_oz[k];
            const int trips = (nk + S - 1) / S;
            for (int t = 0; t < trips; t++)
            {
                const int j = s + t * S;
                bool hit = false;
                if (act && j < nk && !(k == 13 && j == h0 + a))
                {
                    const long js = base + j;
                    float dx = xs - D.xr[js], dy = ys - D.yr[js], dz = zs - D.zr[js];
                    float r2 = dx * dx + dy * dy + dz * dz;
                    hit = r2 < rcut2;
                    if (hit && ag >= 0 && D.grp[js] != ag) hit = false;
                }
                if (hit) hb[fill++][lane] = (uint16_t)((k << 8) | j);
                if (wave_any(fill == HB))
                {
                    for (int h = 0; wave_any(h < fill); h++)
                    {
                        if (h < fill)
                        {
                            unsigned e = hb[h][lane];
                            int kk = e >> 8, jj = e & 255;
                            long js = nb_base[kk] + jj;
                            float dx = (xi - nb_ox[kk]) - D.xr[js];
                            float dy = (yi - nb_oy[kk]) - D.yr[js];
                            float dz = (zi - nb_oz[kk]) - D.zr[js];
                            hit32<PASS>(D, pp, D.nkinds, ki, Fpi, dx, dy, dz, js, ax, ay, az, ae);
                        }
                    }
                    fill = 0;
                }
            }
        }
        for (int h = 0; wave_any(h < fill); h++)
        {
            if (h < fill)
            {
                unsigned e = hb[h][lane];
                int kk = e >> 8, jj = e & 255;
                long js = nb_base[kk] + jj;
                float dx = (xi - nb_ox[kk]) - D.xr[js];
                float dy = (yi - nb_oy[kk]) - D.yr[js];
                float dz = (zi - nb_oz[kk]) - D.zr[js];
                hit32<PASS>(D, pp, D.nkinds, ki, Fpi, dx, dy, dz, js, ax, ay, az, ae);
            }
        }
        __syncthreads();   /* the hit buffer and the reduction scratch share LDS */
        red[0][lane] = (double)ax.s - (double)ax.c;
        red[1][lane] = (double)ay.s - (double)ay.c;
        red[2][lane] = (double)az.s - (double)az.c;
        red[3][lane] = (double)ae.s - (double)ae.c;
        __syncthreads();
        double ecell = 0.0;
        if (lane < nhc)
        {
            double sfx = 0, sfy = 0, sfz = 0, se = 0;
            for (int q = 0; q < S; q++)
            {
                int l = lane + q * nhc;
                sfx += red[0][l]; sfy += red[1][l]; sfz += red[2][l]; se += red[3][l];
            }
            if (PASS == 1)
            {
                double Fp = 0.0;
                if (act)
                {
                    const EamDev &E = D.eam;
                    int el = D.elem[ki];
                    double t = se * E.inv_drho;
                    int k = (int)t;
                    if (k >= E.nrho - 1 || k < 0 || !(se == se))
                    {
                        set_err(D, FMDX_DEV_ERR_RHO_RANGE);
                        k = k < 0 ? 0 : E.nrho - 2;
                    }
                    double b = t - k;
                    const double4 q = E.Fc[(long)el * (E.nrho - 1) + k];
                    Fp = sp_der(q, b, E.inv_drho);
                    ecell = sp_val(q, b);
                }
                D.aux[is] = Fp;
            }
            else
            {
                if (act && T.on)
                {
                    double zabs = (double)zi + oz;
                    int iz = (int)(zabs / T.tcellh_z) + 1;
                    if (iz >= 0 && iz < T.ncz + 2)
                        sfz += T.xi[iz] * D.mass[ki] * (D.vz[is] - T.vcm[3 * iz + 2]);
                }
                D.fx[is] = act ? sfx : 0.0;
                D.fy[is] = act ? sfy : 0.0;
                D.fz[is] = act ? sfz : 0.0;
                ecell = se;
            }
        }
        (void)ox; (void)oy;
        __syncthreads();
        red[0][lane] = ecell;
        __syncthreads();
        if (lane == 0)
            for (int l = 0; l < nhc; l++) cell_acc += red[0][l];
        __syncthreads();
    }
    if (lane == 0)
    {
        if (PASS == 1) D.cell_emb[c] = cell_acc;
        else D.cell_pe[c] = cell_acc;
    }
}

template <int PASS>
__global__ __launch_bounds__(WG) void k_pairs32(KDesc D) { b_pairs32<PASS>(D, blockIdx.x); }

template <int PASS>
__global__ __launch_bounds__(WG) void kb_pairs32(BatchMap M)
{
    const int2 it = M.items[blockIdx.x];
    const KDesc D = M.descs[it.x];
    b_pairs32<PASS>(D, it.y);
}

/* FP32 positions relative to the origin of the home cell, from the FP64 positions */
static __device__ void b_relpos(const KDesc &D, int blk)
{
    long slot = (long)blk * 256 + threadIdx.x;
    if (slot >= (long)D.ncell * D.cap) return;
    int c = (int)(slot / D.cap);
    int k = (int)(slot - (long)c * D.cap);
    if (k >= D.cnt[c]) return;
    int cz = c % D.nc2, t = c / D.nc2, cy = t % D.nc1, cx = t / D.nc1;
    D.xr[slot] = (float)(D.x[slot] - cx * D.h0);
    D.yr[slot] = (float)(D.y[slot] - cy * D.h1);
    D.zr[slot] = (float)(D.z[slot] - cz * D.h2);
}
__global__ __launch_bounds__(256) void k_relpos(KDesc D) { b_relpos(D, blockIdx.x); }
__global__ __launch_bounds__(256) void kb_relpos(BatchMap M)
{
    const int2 it = M.items[blockIdx.x];
    const KDesc D = M.descs[it.x];
    b_relpos(D, it.y);
}

static inline size_t pp_bytes(const KDesc &D) { return sizeof(PairParam) * D.nkinds * D.nkinds; }

/* pass 1 (EAM density + F'(rho)) and pass 0 (forces + energy), launched separately so that
   callers can time each kernel with events on the stream */
hipError_t launch_pass1(const KDesc &D, hipStream_t st, int precision)
{
    if (precision == FMDX_PREC_MIXED)
    {
        unsigned g = (unsigned)(((long)D.ncell * D.cap + 255) / 256);
        hipLaunchKernelGGL(k_relpos, dim3(g), dim3(256), 0, st, D);
        if (D.has_eam) hipLaunchKernelGGL(k_pairs32<1>, dim3(D.ncell), dim3(WG), pp_bytes(D), st, D);
        else launch_fill(D.cell_emb, D.ncell, 0.0, st);
    }
    else
    {
        if (D.has_eam) hipLaunchKernelGGL(k_pairs64<1>, dim3(D.ncell), dim3(WG), pp_bytes(D), st, D);
        else launch_fill(D.cell_emb, D.ncell, 0.0, st);
    }
    return hipGetLastError();
}

hipError_t launch_pass0(const KDesc &D, hipStream_t st, int precision)
{
    if (precision == FMDX_PREC_MIXED) hipLaunchKernelGGL(k_pairs32<0>, dim3(D.ncell), dim3(WG), pp_bytes(D), st, D);
    else hipLaunchKernelGGL(k_pairs64<0>, dim3(D.ncell), dim3(WG), pp_bytes(D), st, D);
    return hipGetLastError();
}

hipError_t launch_forces(const KDesc &D, hipStream_t st, int precision)
{
    hipError_t e = launch_pass1(D, st, precision);
    if (e != hipSuccess) return e;
    e = launch_pass0(D, st, precision);
    if (e != hipSuccess) return e;
    return launch_reduce_pe(D, st);
}

/* every system of a batch: relative positions (mixed), pass 1, pass 0, energies */
hipError_t launchb_forces(const BatchPlan &P, hipStream_t st, int precision)
{
    BatchMap ms = {P.descs, P.slot_items}, mc = {P.descs, P.cell_items};
    const size_t ppb = sizeof(PairParam) * P.max_nkinds * P.max_nkinds;
    if (precision == FMDX_PREC_MIXED)
    {
        hipLaunchKernelGGL(kb_relpos, dim3(P.n_slot), dim3(256), 0, st, ms);
        hipLaunchKernelGGL(kb_pairs32<1>, dim3(P.n_cell), dim3(WG), ppb, st, mc);
        hipLaunchKernelGGL(kb_pairs32<0>, dim3(P.n_cell), dim3(WG), ppb, st, mc);
    }
    else
    {
        hipLaunchKernelGGL(kb_pairs64<1>, dim3(P.n_cell), dim3(WG), ppb, st, mc);
        hipLaunchKernelGGL(kb_pairs64<0>, dim3(P.n_cell), dim3(WG), ppb, st, mc);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launchb_reduce_pe(P, st);
}
