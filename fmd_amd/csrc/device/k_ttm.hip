/*
 * k_ttm.hip -- 1-D two-temperature model on the device.
 *
 *   k_ttm_sum_mv     update_field_vcm_and_mass, turi.c:1112-1178
 *   k_ttm_sum_temp   update_field_temperature_and_number, turi.c:1190-1261
 *   k_ttm_presolve   presolve_common_types_1_2 (+ type2_update_Ce_Ke_G), ttm.c:154-248
 *   k_ttm_solve      ttm_type1_solve_1d / ttm_type2_solve_1d, ttm.c:250-383
 *
 * The atom sums are segmented reductions into turi cells: every block first
 * accumulates into an LDS histogram and then adds its non-empty bins to
 * HBM with one atomic per bin, so the atomic count is about one per block.
 * The FD grid is small (<= FMDX_TTM_MAX_CELLS cells); one workgroup keeps
 * Te, Te2, Ti, Ke, Ce, G and xi in LDS for all sub-steps of an MD step.
 * The activation scan of presolve is sequential in the reference (a cell
 * activated earlier in the scan feeds the next), so one lane runs it.
 *
 * ncx x ncy x ncz turies (no counterpart in the reference, whose 3-D
 * updater is missing, ttm.c:451-452): type 1 with Ce = gamma Te, constant G,
 * and constant conductivity or kappa = K Te / Ti, periodic in x and y.  The
 * atom sums go straight to HBM atomics (the grid is too large for per-block
 * LDS histograms); one workgroup of 1024 lanes keeps Te and Te2 of up to
 * FMDX_TTM_MAX_TCELLS cells in LDS for the sub-steps, each lane owning the
 * same cells throughout so that its xi sums stay in registers.
 */
#include "fmdx_internal.h"

#define TTPB 256



static __device__ __forceinline__ int ttm_bin(const TtmState &S, double z)
{
    return (int)(z / S.tcellh_z) + 1;
}

__global__ __launch_bounds__(TTPB) void k_ttm_clear(TtmState S)
{
    int i = blockIdx.x * TTPB + threadIdx.x;
    if (i >= S.n) return;
    S.mass[i] = 0.0;
    S.vcm[3 * i] = S.vcm[3 * i + 1] = S.vcm[3 * i + 2] = 0.0;
    S.Ti[i] = 0.0;
    S.num[i] = 0;
}

/* sum of m and m v per turi cell (all atoms, no group filter, as in the reference) */
__global__ __launch_bounds__(TTPB) void k_ttm_sum_mv(KDesc D, TtmState S)
{
    __shared__ double hm[FMDX_TTM_MAX_CELLS + 2], hx[FMDX_TTM_MAX_CELLS + 2], hy[FMDX_TTM_MAX_CELLS + 2],
        hz[FMDX_TTM_MAX_CELLS + 2];
    __shared__ int lo, hi;
    for (int i = threadIdx.x; i < S.n; i += TTPB) hm[i] = hx[i] = hy[i] = hz[i] = 0.0;
    if (threadIdx.x == 0) { lo = S.n; hi = -1; }
    __syncthreads();
    long slot = (long)blockIdx.x * TTPB + threadIdx.x;
    if (slot < (long)D.ncell * D.cap)
    {
        int c = (int)(slot / D.cap);
        int k = (int)(slot - (long)c * D.cap);
        if (k < D.cnt[c])
        {
            int iz = ttm_bin(S, D.z[slot]);
            if (iz >= 0 && iz < S.n)
            {
                double m = D.mass[D.kind[slot]];
                atomicAdd(&hm[iz], m);
                atomicAdd(&hx[iz], m * D.vx[slot]);
                atomicAdd(&hy[iz], m * D.vy[slot]);
                atomicAdd(&hz[iz], m * D.vz[slot]);
                atomicMin(&lo, iz);
                atomicMax(&hi, iz);
            }
        }
    }
    __syncthreads();
    for (int i = lo + threadIdx.x; i <= hi; i += TTPB)
        if (hm[i] != 0.0)
        {
            atomicAdd(&S.mass[i], hm[i]);
            atomicAdd(&S.vcm[3 * i], hx[i]);
            atomicAdd(&S.vcm[3 * i + 1], hy[i]);
            atomicAdd(&S.vcm[3 * i + 2], hz[i]);
        }
}

__global__ __launch_bounds__(TTPB) void k_ttm_vcm(TtmState S)
{
    int i = blockIdx.x * TTPB + threadIdx.x;
    if (i < 1 || i > S.ncz) return;
    for (int d = 0; d < 3; d++) S.vcm[3 * i + d] /= S.mass[i];
}

/* sum of m |v - vcm|^2 and atom count per turi cell */
__global__ __launch_bounds__(TTPB) void k_ttm_sum_temp(KDesc D, TtmState S)
{
#pragma clang fp contract(off)
    __shared__ double ht[FMDX_TTM_MAX_CELLS + 2];
    __shared__ unsigned hn[FMDX_TTM_MAX_CELLS + 2];
    __shared__ int lo, hi;
    for (int i = threadIdx.x; i < S.n; i += TTPB) { ht[i] = 0.0; hn[i] = 0; }
    if (threadIdx.x == 0) { lo = S.n; hi = -1; }
    __syncthreads();
    long slot = (long)blockIdx.x * TTPB + threadIdx.x;
    if (slot < (long)D.ncell * D.cap)
    {
        int c = (int)(slot / D.cap);
        int k = (int)(slot - (long)c * D.cap);
        if (k < D.cnt[c])
        {
            int iz = ttm_bin(S, D.z[slot]);
            if (iz >= 0 && iz < S.n)
            {
                double m = D.mass[D.kind[slot]];
                double vx = D.vx[slot] - S.vcm[3 * iz];
                double vy = D.vy[slot] - S.vcm[3 * iz + 1];
                double vz = D.vz[slot] - S.vcm[3 * iz + 2];
                atomicAdd(&ht[iz], m * (vx * vx + vy * vy + vz * vz));
                atomicAdd(&hn[iz], 1u);
                atomicMin(&lo, iz);
                atomicMax(&hi, iz);
            }
        }
    }
    __syncthreads();
    for (int i = lo + threadIdx.x; i <= hi; i += TTPB)
        if (hn[i] != 0)
        {
            atomicAdd(&S.Ti[i], ht[i]);
            atomicAdd(&S.num[i], hn[i]);
        }
}

__global__ __launch_bounds__(TTPB) void k_ttm_temp(TtmState S)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * TTPB + threadIdx.x;
    if (i < 1 || i > S.ncz) return;
    S.Ti[i] /= (3.0 * KB_EV * S.num[i]);
}

/* FIND_klo_AND_ETC (cspline.h:29-44) + SPLINE_VAL (cspline.h:26) */
static __device__ double table_spline(const double *X, const double *Y, const double *YDD, int n, double x)
{
#pragma clang fp contract(off)
    int klo = 0, khi = n - 1;
    while (khi - klo > 1)
    {
        int k = (khi + klo) >> 1;
        if (X[k] > x) khi = k;
        else klo = k;
    }
    double h = X[khi] - X[klo];
    double a = (X[khi] - x) / h;
    double b = (x - X[klo]) / h;
    return (a * Y[klo] + b * Y[khi] + ((a * a * a - a) * YDD[klo] + (b * b * b - b) * YDD[khi]) * (h * h) / 6.0);
}

/* type2_update_Ce_Ke_G, ttm.c:154-192, for one cell */
static __device__ void type2_cell(const TtmState &S, const double *Te, const double *Ti, double *Ke, double *Ce,
                                  double *Gc, int i, unsigned *err)
{
#pragma clang fp contract(off)
    double te = Te[i];
    if (te < 0.) return;
    if (te < S.T_C[0] || te > S.T_C[S.nC - 1]) atomicOr(err, (unsigned)FMDX_DEV_ERR_TTM_CE);
    Ce[i] = table_spline(S.T_C, S.Ct, S.C_DD, S.nC, te);
    Ke[i] = 1. / 3. * Ce[i] * S.Kzh_v2 / (S.Kzh_A * (te * te) + S.Kzh_B * Ti[i]);
    if (te < S.T_G[0] || te > S.T_G[S.nG - 1]) atomicOr(err, (unsigned)FMDX_DEV_ERR_TTM_G);
    Gc[i] = table_spline(S.T_G, S.Gt, S.G_DD, S.nG, te);
}

/* presolve_common_types_1_2 (ttm.c:194-230) and, for type 2, the first Ce/Ke/G update */
__global__ __launch_bounds__(TTPB) void k_ttm_presolve(TtmState S, unsigned *err)
{
#pragma clang fp contract(off)
    if (threadIdx.x == 0)
    {
        for (int i = 1; i <= S.ncz; i++)
        {
            S.xi[i] = 0.0;
            if (S.Te[i] < 0.0 && S.num[i] >= S.min_atoms)
            {
                unsigned count = 0;
                double sum = 0.0;
                int im1 = i - 1;
                if (S.Te[im1] >= 0.0) { sum += S.Te[im1] * S.num[im1]; count += S.num[im1]; }
                if (S.Te[i + 1] >= 0.0) { sum += S.Te[i + 1] * S.num[i + 1]; count += S.num[i + 1]; }
                S.Te[i] = (count > 0) ? sum / (int)count : S.Ti[i];
            }
            if (S.num[i] < S.min_atoms) S.Te[i] = S.Te2[i] = -1.0;
        }
    }
    __syncthreads();
    if (S.type == 2)
        for (int i = 1 + threadIdx.x; i <= S.ncz; i += TTPB) type2_cell(S, S.Te, S.Ti, S.Ke, S.Ce, S.Gc, i, err);
}

/* explicit FD sub-steps of one MD step (ttm.c:250-383) in one workgroup */
__global__ __launch_bounds__(TTPB) void k_ttm_solve(TtmState S, double time, double mddt, unsigned *err)
{
#pragma clang fp contract(off)
    __shared__ double A[FMDX_TTM_MAX_CELLS + 2], B[FMDX_TTM_MAX_CELLS + 2], Ti[FMDX_TTM_MAX_CELLS + 2],
        Ke[FMDX_TTM_MAX_CELLS + 2], Ce[FMDX_TTM_MAX_CELLS + 2], Gc[FMDX_TTM_MAX_CELLS + 2],
        xi[FMDX_TTM_MAX_CELLS + 2];
    for (int i = threadIdx.x; i < S.n; i += TTPB)
    {
        A[i] = S.Te[i];
        B[i] = S.Te2[i];
        Ti[i] = S.Ti[i];
        xi[i] = S.xi[i];
        if (S.type == 2) { Ke[i] = S.Ke[i]; Ce[i] = S.Ce[i]; Gc[i] = S.Gc[i]; }
    }
    __syncthreads();
    const bool calc_source = fabs(time - S.laser_t0) < S.laser_tdiff;
    const double dt = mddt / S.ratio;
    const double dz2 = S.tcellh_z * S.tcellh_z;
    double *Te = A, *Te2 = B;
    for (int j = 0; j < S.ratio; j++)
    {
        double spc = 0.0;
        if (calc_source)
        {
            double q = time + j * dt - S.laser_t0;
            spc = S.laser_factor * exp(S.laser_2sig2_inv * (q * q));
        }
        for (int i = 1 + threadIdx.x; i <= S.ncz; i += TTPB)
        {
            if (Te[i] < 0) continue;
            double source = 0.0;
            if (calc_source)
            {
                double z = (i - 1 + 0.5) * S.tcellh_z;
                source = spc * exp(S.laser_absdepth_inv * (z - S.frontsurf));
            }
            int im1 = i - 1;
            int ilo = (Te[im1] < 0) ? i : im1;
            int ihi = (Te[i + 1] < 0) ? i : i + 1;
            if (S.type == 1)
                Te2[i] = Te[i] + dt / (S.C_gamma * Te[i]) *
                                     (S.K * (Te[ihi] - 2 * Te[i] + Te[ilo]) / dz2 - S.G * (Te[i] - Ti[i]) + source);
            else
            {
                double KeL = 0.5 * (Ke[i] + Ke[ilo]);
                double KeR = 0.5 * (Ke[i] + Ke[ihi]);
                Te2[i] = Te[i] + dt / Ce[i] *
                                     ((KeR * (Te[ihi] - Te[i]) - KeL * (Te[i] - Te[ilo])) / dz2 -
                                      Gc[i] * (Te[i] - Ti[i]) + source);
            }
            xi[i] += Te2[i];
        }
        __syncthreads();
        double *t = Te2; Te2 = Te; Te = t;
        if (S.type == 2 && j < S.ratio - 1)
        {
            for (int i = 1 + threadIdx.x; i <= S.ncz; i += TTPB) type2_cell(S, Te, Ti, Ke, Ce, Gc, i, err);
            __syncthreads();
        }
    }
    for (int i = 1 + threadIdx.x; i <= S.ncz; i += TTPB)
        if (S.num[i] >= S.min_atoms)
        {
            double G = S.type == 1 ? S.G : Gc[i];
            xi[i] = G * S.tvol * (xi[i] / S.ratio - Ti[i]) / (3 * KB_EV * S.num[i] * Ti[i]);
        }
    __syncthreads();
    for (int i = threadIdx.x; i < S.n; i += TTPB)
    {
        S.Te[i] = Te[i];
        S.Te2[i] = Te2[i];
        S.xi[i] = xi[i];
        if (S.type == 2) { S.Ke[i] = Ke[i]; S.Ce[i] = Ce[i]; S.Gc[i] = Gc[i]; }
    }
}

/* ------------------------------------------------ ncx x ncy x ncz turies */

/* turi cell of a position, x and y clamped to the grid; a non-finite or
   far-away z goes to the lower ghost cell (the comparisons run in FP64 so
   that no out-of-range conversion happens) */
static __device__ __forceinline__ int tcell3(const TtmState &S, double x, double y, double z)
{
    const double fx = x / S.tcellh_x, fy = y / S.tcellh_y, fz = z / S.tcellh_z;
    const int ix = fx >= 0.0 ? (fx < S.ncx ? (int)fx : S.ncx - 1) : 0;
    const int iy = fy >= 0.0 ? (fy < S.ncy ? (int)fy : S.ncy - 1) : 0;
    const int iz = (fz >= -1.0 && fz < S.nzg - 1) ? (int)fz + 1 : 0;
    return (ix * S.ncy + iy) * S.nzg + iz;
}

__global__ __launch_bounds__(TPB) void k_ttm_sum_mv3(KDesc D, TtmState S)
{
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        (void)c;
        const int a = tcell3(S, D.x[slot], D.y[slot], D.z[slot]);
        if (a < 0) return;
        const double m = D.mass[D.kind[slot]];
        atomicAdd(&S.mass[a], m);
        atomicAdd(&S.vcm[3 * a], m * D.vx[slot]);
        atomicAdd(&S.vcm[3 * a + 1], m * D.vy[slot]);
        atomicAdd(&S.vcm[3 * a + 2], m * D.vz[slot]);
    });
}

__global__ __launch_bounds__(TTPB) void k_ttm_vcm3(TtmState S)
{
    const int a = blockIdx.x * TTPB + threadIdx.x;
    if (a >= S.n || S.mass[a] == 0.0) return;
    for (int d = 0; d < 3; d++) S.vcm[3 * a + d] /= S.mass[a];
}

__global__ __launch_bounds__(TPB) void k_ttm_sum_temp3(KDesc D, TtmState S)
{
#pragma clang fp contract(off)
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        (void)c;
        const int a = tcell3(S, D.x[slot], D.y[slot], D.z[slot]);
        if (a < 0) return;
        const double m = D.mass[D.kind[slot]];
        const double vx = D.vx[slot] - S.vcm[3 * a];
        const double vy = D.vy[slot] - S.vcm[3 * a + 1];
        const double vz = D.vz[slot] - S.vcm[3 * a + 2];
        atomicAdd(&S.Ti[a], m * (vx * vx + vy * vy + vz * vz));
        atomicAdd(&S.num[a], 1u);
    });
}

__global__ __launch_bounds__(TTPB) void k_ttm_temp3(TtmState S)
{
#pragma clang fp contract(off)
    const int a = blockIdx.x * TTPB + threadIdx.x;
    if (a >= S.n || S.num[a] == 0) return;
    S.Ti[a] /= (3.0 * KB_EV * S.num[a]);
}

/* the activation scan of presolve_common_types_1_2 over the grid, cells in
   index order; a newly active cell takes the atom-weighted mean Te of its
   active z, x and y neighbours (z first, as the 1-D scan) */
__global__ __launch_bounds__(64) void k_ttm_presolve3(TtmState S)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    for (int ix = 0; ix < S.ncx; ix++)
        for (int iy = 0; iy < S.ncy; iy++)
            for (int iz = 1; iz <= S.ncz; iz++)
            {
                const int a = (ix * S.ncy + iy) * S.nzg + iz;
                S.xi[a] = 0.0;
                if (S.Te[a] < 0.0 && S.num[a] >= S.min_atoms)
                {
                    const int nb[6] = {a - 1, a + 1,
                                       (((ix + S.ncx - 1) % S.ncx) * S.ncy + iy) * S.nzg + iz,
                                       (((ix + 1) % S.ncx) * S.ncy + iy) * S.nzg + iz,
                                       (ix * S.ncy + (iy + S.ncy - 1) % S.ncy) * S.nzg + iz,
                                       (ix * S.ncy + (iy + 1) % S.ncy) * S.nzg + iz};
                    unsigned count = 0;
                    double sum = 0.0;
                    for (int k = 0; k < 6; k++)
                        if (S.Te[nb[k]] >= 0.0)
                        {
                            sum += S.Te[nb[k]] * S.num[nb[k]];
                            count += S.num[nb[k]];
                        }
                    S.Te[a] = (count > 0) ? sum / (int)count : S.Ti[a];
                }
                if (S.num[a] < S.min_atoms) S.Te[a] = S.Te2[a] = -1.0;
            }
}

#define T3PB 1024
#define T3OWN 4   /* cells owned per lane: T3PB * T3OWN >= FMDX_TTM_MAX_TCELLS */

/* explicit FD sub-steps on the grid: Ce = C_gamma Te, constant G; kmode 0:
   K (Te_hi - 2 Te + Te_lo) / h^2 per dimension (the 1-D type-1 expression,
   ttm.c:281-287), kmode 1: kappa = K Te / Ti with face values
   (kappa_i + kappa_nb) / 2 (the type-2 flux form, ttm.c:347-352); an inactive
   neighbour is replaced by the cell itself */
__global__ __launch_bounds__(T3PB) void k_ttm_solve3(TtmState S, double time, double mddt)
{
#pragma clang fp contract(off)
    extern __shared__ double t3sh[];
    double *Te = t3sh, *Te2 = t3sh + S.n;
    for (int a = threadIdx.x; a < S.n; a += T3PB)
    {
        Te[a] = S.Te[a];
        Te2[a] = S.Te2[a];
    }
    __syncthreads();
    double xacc[T3OWN];
    for (int k = 0; k < T3OWN; k++) xacc[k] = 0.0;
    const bool calc_source = fabs(time - S.laser_t0) < S.laser_tdiff;
    const double dt = mddt / S.ratio;
    const double h2[3] = {S.tcellh_x * S.tcellh_x, S.tcellh_y * S.tcellh_y, S.tcellh_z * S.tcellh_z};
    for (int j = 0; j < S.ratio; j++)
    {
        double spc = 0.0;
        if (calc_source)
        {
            const double q = time + j * dt - S.laser_t0;
            spc = S.laser_factor * exp(S.laser_2sig2_inv * (q * q));
        }
        for (int k = 0; k < T3OWN; k++)
        {
            const int a = threadIdx.x + k * T3PB;
            if (a >= S.n) break;
            const int iz = a % S.nzg, col = a / S.nzg;
            if (iz < 1 || iz > S.ncz || Te[a] < 0) continue;
            const int ix = col / S.ncy, iy = col % S.ncy;
            double source = 0.0;
            if (calc_source)
            {
                const double z = (iz - 1 + 0.5) * S.tcellh_z;
                source = spc * exp(S.laser_absdepth_inv * (z - S.frontsurf));
            }
            int lo[3] = {(((ix + S.ncx - 1) % S.ncx) * S.ncy + iy) * S.nzg + iz,
                         (ix * S.ncy + (iy + S.ncy - 1) % S.ncy) * S.nzg + iz, a - 1};
            int hi[3] = {(((ix + 1) % S.ncx) * S.ncy + iy) * S.nzg + iz,
                         (ix * S.ncy + (iy + 1) % S.ncy) * S.nzg + iz, a + 1};
            for (int d = 0; d < 3; d++)
            {
                if (Te[lo[d]] < 0) lo[d] = a;
                if (Te[hi[d]] < 0) hi[d] = a;
            }
            const double Ti = S.Ti[a];
            double flux = 0.0;
            if (S.kmode == 0)
            {
                for (int d = 0; d < 3; d++) flux += S.K * (Te[hi[d]] - 2 * Te[a] + Te[lo[d]]) / h2[d];
            }
            else
            {
                const double ka = S.K * Te[a] / Ti;
                for (int d = 0; d < 3; d++)
                {
                    const double KeL = 0.5 * (ka + S.K * Te[lo[d]] / S.Ti[lo[d]]);
                    const double KeR = 0.5 * (ka + S.K * Te[hi[d]] / S.Ti[hi[d]]);
                    flux += (KeR * (Te[hi[d]] - Te[a]) - KeL * (Te[a] - Te[lo[d]])) / h2[d];
                }
            }
            Te2[a] = Te[a] + dt / (S.C_gamma * Te[a]) * (flux - S.G * (Te[a] - Ti) + source);
            xacc[k] += Te2[a];
        }
        __syncthreads();
        double *t = Te2; Te2 = Te; Te = t;
    }
    for (int k = 0; k < T3OWN; k++)
    {
        const int a = threadIdx.x + k * T3PB;
        if (a >= S.n) break;
        const int iz = a % S.nzg;
        if (iz >= 1 && iz <= S.ncz)
            S.xi[a] = S.num[a] >= S.min_atoms
                          ? S.G * S.tvol * (xacc[k] / S.ratio - S.Ti[a]) / (3 * KB_EV * S.num[a] * S.Ti[a])
                          : 0.0;
        S.Te[a] = Te[a];
        S.Te2[a] = Te2[a];
    }
}

hipError_t launch_ttm_fields(const KDesc &D, const TtmState &S, hipStream_t st)
{
    const unsigned gn = (S.n + TTPB - 1) / TTPB;
    hipLaunchKernelGGL(k_ttm_clear, dim3(gn), dim3(TTPB), 0, st, S);
    if (S.dim3)
    {
        const unsigned gc = (unsigned)((D.ncell + CPB - 1) / CPB);
        hipLaunchKernelGGL(k_ttm_sum_mv3, dim3(gc), dim3(TPB), 0, st, D, S);
        hipLaunchKernelGGL(k_ttm_vcm3, dim3(gn), dim3(TTPB), 0, st, S);
        hipLaunchKernelGGL(k_ttm_sum_temp3, dim3(gc), dim3(TPB), 0, st, D, S);
        hipLaunchKernelGGL(k_ttm_temp3, dim3(gn), dim3(TTPB), 0, st, S);
        return hipGetLastError();
    }
    const unsigned gs = (unsigned)(((long)D.ncell * D.cap + TTPB - 1) / TTPB);
    hipLaunchKernelGGL(k_ttm_sum_mv, dim3(gs), dim3(TTPB), 0, st, D, S);
    hipLaunchKernelGGL(k_ttm_vcm, dim3(gn), dim3(TTPB), 0, st, S);
    hipLaunchKernelGGL(k_ttm_sum_temp, dim3(gs), dim3(TTPB), 0, st, D, S);
    hipLaunchKernelGGL(k_ttm_temp, dim3(gn), dim3(TTPB), 0, st, S);
    return hipGetLastError();
}

hipError_t launch_ttm_presolve(const TtmState &S, unsigned *err, hipStream_t st)
{
    if (S.dim3) hipLaunchKernelGGL(k_ttm_presolve3, dim3(1), dim3(64), 0, st, S);
    else hipLaunchKernelGGL(k_ttm_presolve, dim3(1), dim3(TTPB), 0, st, S, err);
    return hipGetLastError();
}

hipError_t launch_ttm_solve(const TtmState &S, double time, double dt, unsigned *err, hipStream_t st)
{
    if (S.dim3)
        hipLaunchKernelGGL(k_ttm_solve3, dim3(1), dim3(T3PB), sizeof(double) * 2 * S.n, st, S, time, dt);
    else
        hipLaunchKernelGGL(k_ttm_solve, dim3(1), dim3(TTPB), 0, st, S, time, dt, err);
    return hipGetLastError();
}
