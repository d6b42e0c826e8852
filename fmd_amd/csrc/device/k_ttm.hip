/*
 * k_ttm.hip -- two-temperature model on the device.
 *
 *   k_ttm_sums       update_field_vcm_and_mass + update_field_temperature_and_number,
 *                    turi.c:1112-1261
 *   k_ttm_presolve   presolve_common_types_1_2 (+ type2_update_Ce_Ke_G), ttm.c:154-248
 *   k_ttm_solve      ttm_type1_solve_1d / ttm_type2_solve_1d, ttm.c:250-383
 *
 * The atom sums are gathers, not scatters: one workgroup per turi cell walks
 * the MD cells that can hold its atoms (the cells its box overlaps, with a
 * margin of one for rounding), keeps the atoms whose turi index is its own,
 * sums them lane by lane and combines the lanes in a fixed tree, first m and
 * m v (then vcm), then m |v - vcm|^2 and the count (then Ti).  No atomics:
 * every run gives the same bits.  The FD grid is small (<= FMDX_TTM_MAX_CELLS
 * cells); one workgroup keeps Te, Te2, Ti, Ke, Ce, G and xi in LDS for all
 * sub-steps of an MD step.  The activation scan of presolve is sequential in
 * the reference (a cell activated earlier in the scan feeds the next), so one
 * lane runs it.
 *
 * ncx x ncy x ncz turies (no counterpart in the reference, whose 3-D
 * updater is missing, ttm.c:451-452): type 1 with Ce = gamma Te, constant G,
 * and constant conductivity or kappa = K Te / Ti, periodic in x and y.  The
 * presolve decides every cell at once and then runs one lane per (x, y) column
 * through its activations in z order, as the 1-D scan takes them;
 * one workgroup of 1024 lanes keeps Te (and the cells' kappa) of up to
 * FMDX_TTM_MAX_TCELLS cells in LDS for the sub-steps, each lane owning the
 * same cells throughout so that its new values and xi sums stay in registers.
 */
#include "fmdx_internal.h"

#define TTPB 256
#define T3PB_P 256   /* lanes of the grid presolve (one per column) */




/* turi cell of a position on an ncx x ncy x ncz turi: x and y clamped to the
   grid; a non-finite or far-away z goes to the lower ghost cell (the
   comparisons run in FP64 so that no out-of-range conversion happens) */
static __device__ __forceinline__ int tcell3(const TtmState &S, double x, double y, double z)
{
    const double fx = x / S.tcellh_x, fy = y / S.tcellh_y, fz = z / S.tcellh_z;
    const int ix = fx >= 0.0 ? (fx < S.ncx ? (int)fx : S.ncx - 1) : 0;
    const int iy = fy >= 0.0 ? (fy < S.ncy ? (int)fy : S.ncy - 1) : 0;
    const int iz = (fz >= -1.0 && fz < S.nzg - 1) ? (int)fz + 1 : 0;
    return (ix * S.ncy + iy) * S.nzg + iz;
}

/* MD cells [lo, hi] along one dimension that can hold atoms of turi slab i of
   n (width tw; MD cells of width h, nc of them): the overlap plus one cell of
   margin; the first and last slabs also take everything beyond them */
static __device__ __forceinline__ void md_range(int i, int n, double tw, double h, int nc, int &lo, int &hi)
{
    lo = i <= 0 ? 0 : (int)floor(i * tw / h) - 1;
    hi = i >= n - 1 ? nc - 1 : (int)floor((i + 1) * tw / h) + 1;
    lo = lo < 0 ? 0 : (lo > nc - 1 ? nc - 1 : lo);
    hi = hi < lo ? lo : (hi > nc - 1 ? nc - 1 : hi);
}

/* sums of lane values in a fixed order (tree over the workgroup), N values at once
   (sh holds N x TTPB) with one barrier per tree level; every lane gets the results */
template <int N>
static __device__ __forceinline__ void block_sums(double (&v)[N], double *sh)
{
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; i++) sh[i * TTPB + threadIdx.x] = v[i];
    __syncthreads();
    for (int o = TTPB / 2; o > 0; o >>= 1)
    {
        if ((int)threadIdx.x < o)
        {
#pragma unroll
            for (int i = 0; i < N; i++) sh[i * TTPB + threadIdx.x] += sh[i * TTPB + threadIdx.x + o];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = sh[i * TTPB];
}

/* turi index of every atom, into the binning scratch word mv (free between
   steps: the next re-binning rewrites it for every atom before reading it).
   1-D turi (G3 false): (int)(z / tcellh_z) + 1 as turi.c:1136, -1 outside
   [0, n); grid turi: tcell3. */
template <bool G3>
__global__ __launch_bounds__(TPB) void k_ttm_index(KDesc D, TtmState S)
{
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        (void)c;
        int a;
        if (G3) a = tcell3(S, D.x[slot], D.y[slot], D.z[slot]);
        else
        {
            a = (int)(D.z[slot] / S.tcellh_z) + 1;
            a = (a >= 0 && a < S.n) ? a : -1;
        }
        D.mv[slot] = a;
    });
}

/* Per turi cell a (one workgroup): mass, vcm, Ti and the atom count of the
   atoms whose turi index is a (all atoms, no group filter, as in the
   reference), from the MD cells that can hold them, 16 lanes per MD cell.
   1-D turi (G3 false): as turi.c:1146-1150, 1244-1248 vcm and Ti are divided
   in every owned cell 1..ncz, empty or not, and the ghost cells keep the raw
   sums.  Grid turi (G3): only non-empty cells are divided. */
template <bool G3>
__global__ __launch_bounds__(TTPB) void k_ttm_sums(KDesc D, TtmState S)
{
#pragma clang fp contract(off)
    __shared__ double sh[4 * TTPB];
    const int a = blockIdx.x;
    const int iz = a % S.nzg, col = a / S.nzg;
    int x0 = 0, x1 = D.nc0 - 1, y0 = 0, y1 = D.nc1 - 1, z0, z1;
    if (G3)
    {
        md_range(col / S.ncy, S.ncx, S.tcellh_x, D.h0, D.nc0, x0, x1);
        md_range(col % S.ncy, S.ncy, S.tcellh_y, D.h1, D.nc1, y0, y1);
    }
    md_range(iz - 1, S.ncz, S.tcellh_z, D.h2, D.nc2, z0, z1);
    if (iz == 0) { z0 = 0; z1 = D.nc2 > 1 ? 1 : 0; }
    if (iz == S.nzg - 1) { z1 = D.nc2 - 1; }
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
    const int items = nx * ny * nz * LPC;
    /* item k: MD cell k / LPC of the box, atoms k % LPC, + LPC, ... of it.  A lane
       walks its items in order and adds its atoms in that order in both sums.  The
       walk is latency-bound (per item a chain cnt -> mv -> kind, v), so the counts
       and first turi indices of U items load together, their first atoms' kind and
       velocity load together, and the first sum keeps the slots it takes (KMAX per
       lane, in walking order) for the second sum to replay instead of walking again;
       a lane with more resumes its walk at its (KMAX+1)-th slot.  Slots are kept as
       32 bits: a slot holds more than 64 bytes of state, so 2^32 of them exceed HBM. */
    constexpr int U = 4, KMAX = 16;   /* 8: 145 VGPRs, 3 waves per SIMD, slower */
    const int j0 = (int)threadIdx.x % LPC;   /* TTPB is a multiple of LPC: each item starts at j0 */
    auto cell_of = [&](int k) __attribute__((always_inline)) {
        const int b = k / LPC, cz = z0 + b % nz, t = b / nz;
        return ((x0 + t / ny) * D.nc1 + (y0 + t % ny)) * D.nc2 + cz;
    };
    unsigned hit[KMAX];
    int nh = 0, kres = items, jres = 0;
    double m = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    auto take = [&](long s, int kd, double vx, double vy, double vz, int k, int j) __attribute__((always_inline)) {
        const double mm = D.mass[kd];
        m += mm;
        px += mm * vx;
        py += mm * vy;
        pz += mm * vz;
#pragma unroll
        for (int q = 0; q < KMAX; q++)
            if (q == nh) hit[q] = (unsigned)s;
        if (nh == KMAX)
        {
            kres = k;
            jres = j;
        }
        nh++;
    };
    for (int k0 = threadIdx.x; k0 < items; k0 += U * TTPB)
    {
        int n_[U], w_[U], kd_[U];
        long b_[U];
        double vx_[U], vy_[U], vz_[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const int k = k0 + u * TTPB;
            n_[u] = 0;
            w_[u] = -1;
            b_[u] = 0;
            if (k < items)
            {
                const int c = cell_of(k);
                b_[u] = (long)c * D.cap;
                n_[u] = D.cnt[c];
                if (j0 < D.cap) w_[u] = D.mv[b_[u] + j0];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            kd_[u] = 0;
            vx_[u] = vy_[u] = vz_[u] = 0.0;
            if (j0 < n_[u] && w_[u] == a)
            {
                const long s = b_[u] + j0;
                kd_[u] = D.kind[s];
                vx_[u] = D.vx[s];
                vy_[u] = D.vy[s];
                vz_[u] = D.vz[s];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const int k = k0 + u * TTPB;
            if (j0 < n_[u] && w_[u] == a) take(b_[u] + j0, kd_[u], vx_[u], vy_[u], vz_[u], k, j0);
            for (int j = j0 + LPC; j < n_[u]; j += LPC)
            {
                const long s = b_[u] + j;
                if (D.mv[s] == a) take(s, D.kind[s], D.vx[s], D.vy[s], D.vz[s], k, j);
            }
        }
    }
    {
        double v4[4] = {m, px, py, pz};
        block_sums(v4, sh);
        m = v4[0];
        px = v4[1];
        py = v4[2];
        pz = v4[3];
    }
    const bool owned = iz >= 1 && iz <= S.ncz;
    const bool divide = G3 ? m != 0.0 : owned;
    const double vx0 = divide ? px / m : px, vy0 = divide ? py / m : py, vz0 = divide ? pz / m : pz;
    double e = 0.0, cnt = 0.0;
    auto add2 = [&](int kd, double vxs, double vys, double vzs) __attribute__((always_inline)) {
        const double mm = D.mass[kd];
        const double vx = vxs - vx0, vy = vys - vy0, vz = vzs - vz0;
        e += mm * (vx * vx + vy * vy + vz * vz);
        cnt += 1.0;
    };
    /* the kept slots, eight loads in flight at a time */
#pragma unroll
    for (int q0 = 0; q0 < KMAX; q0 += 8)
    {
        int kd[8];
        double ux[8], uy[8], uz[8];
#pragma unroll
        for (int q = 0; q < 8; q++)
        {
            kd[q] = 0;
            ux[q] = uy[q] = uz[q] = 0.0;
            if (q0 + q < nh)
            {
                const long s = hit[q0 + q];
                kd[q] = D.kind[s];
                ux[q] = D.vx[s];
                uy[q] = D.vy[s];
                uz[q] = D.vz[s];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q0 + q < nh) add2(kd[q], ux[q], uy[q], uz[q]);
    }
    /* a lane with more than KMAX slots walks on from its (KMAX+1)-th */
    for (int k = kres; k < items; k += TTPB)
    {
        const int c = cell_of(k);
        const long base = (long)c * D.cap;
        for (int j = k == kres ? jres : j0, n = D.cnt[c]; j < n; j += LPC)
        {
            const long s = base + j;
            if (D.mv[s] == a) add2(D.kind[s], D.vx[s], D.vy[s], D.vz[s]);
        }
    }
    {
        double v2[2] = {e, cnt};
        block_sums(v2, sh);
        e = v2[0];
        cnt = v2[1];
    }
    if (threadIdx.x == 0)
    {
        const unsigned nn = (unsigned)cnt;
        S.mass[a] = m;
        S.vcm[3 * a] = vx0;
        S.vcm[3 * a + 1] = vy0;
        S.vcm[3 * a + 2] = vz0;
        S.num[a] = nn;
        S.Ti[a] = (G3 ? nn != 0 : owned) ? e / (3.0 * KB_EV * nn) : e;
    }
}

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

/* Activation / deactivation on the grid (the extension of
   presolve_common_types_1_2, ttm.c:194-230): each column is scanned along z
   upwards as the 1-D scan does (the cell below already processed in this pass,
   the cell above not yet); the x and y neighbours contribute their Te from
   before the pass, kept in LDS, so the columns are independent.  A newly
   active cell takes the atom-weighted mean Te of its active neighbours in the
   order z-, z+, x-, x+, y-, y+, or its Ti when it has none.
   Whether a cell activates (Te < 0 with enough atoms) or is switched off (too
   few atoms) rests on its own Te before the pass and its atom count, so all
   lanes decide every cell and write the switch-offs and the xi reset at once;
   one lane per column then walks only the activations in z order, each reading
   the cell below as the scan left it and the cell above as it was before. */
__global__ __launch_bounds__(T3PB_P) void k_ttm_presolve3(TtmState S)
{
#pragma clang fp contract(off)
    extern __shared__ double before[];   /* S.n values of Te, then S.n activation flags */
    unsigned char *act = (unsigned char *)(before + S.n);
    for (int a = threadIdx.x; a < S.n; a += T3PB_P)
    {
        const double te = S.Te[a];
        before[a] = te;
        const int iz = a % S.nzg;
        bool on = false;
        if (iz >= 1 && iz <= S.ncz)
        {
            S.xi[a] = 0.0;
            on = te < 0.0 && S.num[a] >= S.min_atoms;
            if (S.num[a] < S.min_atoms) S.Te[a] = S.Te2[a] = -1.0;
        }
        act[a] = on;
    }
    __syncthreads();   /* the switch-offs above are what the activations below read */
    for (int col = threadIdx.x; col < S.ncx * S.ncy; col += T3PB_P)
    {
        const int ix = col / S.ncy, iy = col % S.ncy;
        const int xl = ((ix + S.ncx - 1) % S.ncx) * S.ncy + iy, xh = ((ix + 1) % S.ncx) * S.ncy + iy;
        const int yl = ix * S.ncy + (iy + S.ncy - 1) % S.ncy, yh = ix * S.ncy + (iy + 1) % S.ncy;
        for (int iz = 1; iz <= S.ncz; iz++)
        {
            const int a = col * S.nzg + iz;
            if (!act[a]) continue;
            const int nb[6] = {a - 1, a + 1, xl * S.nzg + iz, xh * S.nzg + iz, yl * S.nzg + iz, yh * S.nzg + iz};
            unsigned count = 0;
            double sum = 0.0;
            for (int k = 0; k < 6; k++)
            {
                const double te = k == 0 ? S.Te[nb[k]] : before[nb[k]];
                if (te >= 0.0)
                {
                    sum += te * S.num[nb[k]];
                    count += S.num[nb[k]];
                }
            }
            S.Te[a] = (count > 0) ? sum / (int)count : S.Ti[a];
        }
    }
}

#define T3PB 1024
#define T3OWN 4   /* cells owned per lane: T3PB * T3OWN >= FMDX_TTM_MAX_TCELLS */

/* explicit FD sub-steps on the grid: Ce = C_gamma Te, constant G; kmode 0:
   K (Te_hi - 2 Te + Te_lo) / h^2 per dimension (the 1-D type-1 expression,
   ttm.c:281-287), kmode 1: kappa = K Te / Ti with face values
   (kappa_i + kappa_nb) / 2 (the type-2 flux form, ttm.c:347-352); an inactive
   neighbour is replaced by the cell itself.  Each lane owns up to T3OWN cells
   for the whole call: their new Te, the sum of their sub-step values (for xi)
   and the depth factor of the laser source, exp((z - front) / -Lp), which
   does not change between sub-steps, stay in registers.  Per sub-step the
   cells' kappa are computed once into LDS (each is used by seven cells), the
   new values are computed from LDS, and written back after a barrier. */
__global__ __launch_bounds__(T3PB) void k_ttm_solve3(TtmState S, double time, double mddt)
{
#pragma clang fp contract(off)
    extern __shared__ double t3sh[];
    double *Te = t3sh, *kap = t3sh + S.n;
    for (int a = threadIdx.x; a < S.n; a += T3PB) Te[a] = S.Te[a];
    const bool calc_source = fabs(time - S.laser_t0) < S.laser_tdiff;
    const double dt = mddt / S.ratio;
    /* the divisors that do not change during the call are inverted once: 1 / h^2 per
       dimension and K / Ti per owned cell (one FP64 division per cell and sub-step is left,
       dt / (C Te); the single workgroup's sub-steps are bound by its VALU) */
    const double ih2[3] = {1.0 / (S.tcellh_x * S.tcellh_x), 1.0 / (S.tcellh_y * S.tcellh_y),
                           1.0 / (S.tcellh_z * S.tcellh_z)};
    /* other: the cell's value in the second buffer of the reference's swap; a
       cell that is not updated in a sub-step takes it, as a swap would give.
       The owned cells' neighbours (x-, y-, z-, x+, y+, z+), Ti and whether they
       are interior are fixed for the call. */
    double xacc[T3OWN], nt[T3OWN], other[T3OWN], depth[T3OWN], tiv[T3OWN], kti[T3OWN];
    int nbi[T3OWN][6];
    bool inner[T3OWN];
    for (int k = 0; k < T3OWN; k++)
    {
        xacc[k] = 0.0;
        depth[k] = 0.0;
        tiv[k] = 0.0;
        kti[k] = 0.0;
        inner[k] = false;
        const int a = threadIdx.x + k * T3PB;
        for (int d = 0; d < 6; d++) nbi[k][d] = a;
        other[k] = a < S.n ? S.Te2[a] : 0.0;
        if (a >= S.n) continue;
        const int iz = a % S.nzg, col = a / S.nzg;
        const int ix = col / S.ncy, iy = col % S.ncy;
        inner[k] = iz >= 1 && iz <= S.ncz;
        tiv[k] = S.Ti[a];
        kti[k] = S.K / tiv[k];
        if (inner[k])
        {
            nbi[k][0] = (((ix + S.ncx - 1) % S.ncx) * S.ncy + iy) * S.nzg + iz;
            nbi[k][1] = (ix * S.ncy + (iy + S.ncy - 1) % S.ncy) * S.nzg + iz;
            nbi[k][2] = a - 1;
            nbi[k][3] = (((ix + 1) % S.ncx) * S.ncy + iy) * S.nzg + iz;
            nbi[k][4] = (ix * S.ncy + (iy + 1) % S.ncy) * S.nzg + iz;
            nbi[k][5] = a + 1;
        }
        if (calc_source)
        {
            const double z = (iz - 1 + 0.5) * S.tcellh_z;
            depth[k] = exp(S.laser_absdepth_inv * (z - S.frontsurf));
        }
    }
    /* the source's time factor of sub-step j depends on j alone: with up to 64
       sub-steps, lane j of every wave evaluates it once and sub-step j reads it
       from that lane */
    const bool spc_lanes = calc_source && S.ratio <= 64;
    double spc_l = 0.0;
    if (spc_lanes && (int)(threadIdx.x & 63) < S.ratio)
    {
        const int j = threadIdx.x & 63;
        const double q = time + j * dt - S.laser_t0;
        spc_l = S.laser_factor * exp(S.laser_2sig2_inv * (q * q));
    }
    __syncthreads();
    for (int j = 0; j < S.ratio; j++)
    {
        double spc = 0.0;
        if (spc_lanes) spc = __shfl(spc_l, j, 64);
        else if (calc_source)
        {
            const double q = time + j * dt - S.laser_t0;
            spc = S.laser_factor * exp(S.laser_2sig2_inv * (q * q));
        }
        if (S.kmode == 1)
        {
            for (int k = 0; k < T3OWN; k++)
            {
                const int a = threadIdx.x + k * T3PB;
                if (a < S.n) kap[a] = kti[k] * Te[a];
            }
            __syncthreads();
        }
        for (int k = 0; k < T3OWN; k++)
        {
            const int a = threadIdx.x + k * T3PB;
            nt[k] = other[k];
            if (a >= S.n || !inner[k] || Te[a] < 0) continue;
            int lo[3] = {nbi[k][0], nbi[k][1], nbi[k][2]};
            int hi[3] = {nbi[k][3], nbi[k][4], nbi[k][5]};
            for (int d = 0; d < 3; d++)
            {
                if (Te[lo[d]] < 0) lo[d] = a;
                if (Te[hi[d]] < 0) hi[d] = a;
            }
            const double Ti = tiv[k];
            double flux = 0.0;
            if (S.kmode == 0)
            {
                for (int d = 0; d < 3; d++) flux += S.K * (Te[hi[d]] - 2 * Te[a] + Te[lo[d]]) * ih2[d];
            }
            else
            {
                const double ka = kap[a];
                for (int d = 0; d < 3; d++)
                {
                    const double KeL = 0.5 * (ka + kap[lo[d]]);
                    const double KeR = 0.5 * (ka + kap[hi[d]]);
                    flux += (KeR * (Te[hi[d]] - Te[a]) - KeL * (Te[a] - Te[lo[d]])) * ih2[d];
                }
            }
            const double source = calc_source ? spc * depth[k] : 0.0;
            nt[k] = Te[a] + dt / (S.C_gamma * Te[a]) * (flux - S.G * (Te[a] - Ti) + source);
            xacc[k] += nt[k];
        }
        __syncthreads();
        for (int k = 0; k < T3OWN; k++)
        {
            const int a = threadIdx.x + k * T3PB;
            if (a >= S.n) continue;
            other[k] = Te[a];
            Te[a] = nt[k];
        }
        __syncthreads();
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
        S.Te2[a] = other[k];
    }
}

hipError_t launch_ttm_fields(const KDesc &D, const TtmState &S, hipStream_t st)
{
    const dim3 gs((D.ncell + CPB - 1) / CPB);
    if (S.dim3)
    {
        hipLaunchKernelGGL(k_ttm_index<true>, gs, dim3(TPB), 0, st, D, S);
        hipLaunchKernelGGL(k_ttm_sums<true>, dim3(S.n), dim3(TTPB), 0, st, D, S);
    }
    else
    {
        hipLaunchKernelGGL(k_ttm_index<false>, gs, dim3(TPB), 0, st, D, S);
        hipLaunchKernelGGL(k_ttm_sums<false>, dim3(S.n), dim3(TTPB), 0, st, D, S);
    }
    return hipGetLastError();
}

hipError_t launch_ttm_presolve(const TtmState &S, unsigned *err, hipStream_t st)
{
    if (S.dim3) hipLaunchKernelGGL(k_ttm_presolve3, dim3(1), dim3(T3PB_P), (sizeof(double) + 1) * S.n, st, S);
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
