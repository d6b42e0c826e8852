/*
 * k_step.hip -- O(N) step kernels: velocity Verlet / Berendsen / symplectic
 * Euler, the exact-order cell re-binning, and the deterministic reductions.
 *
 * Integration arithmetic follows the reference expression by expression with
 * FMA contraction disabled, so one FP64 step from identical inputs gives the
 * reference's bits (integrators.c:63-70, 115-121, 144-145).
 *
 * Re-binning reproduces _fmd_refreshGrid + _fmd_particles_migrate
 * (misc.c:268-308, md-ghost.c:827-869) without their sequential walk:
 *   1. every atom finds its destination (fused into the integrator kernels);
 *      arrivals per destination are counted, and "early" arrivals -- direct
 *      moves from a cell of lower storage index, which the reference appends
 *      before the destination cell is itself scanned -- are counted apart;
 *   2. per cell with leavers, the reference's swap-with-last scan
 *      (cell.c:151-160) is replayed on indices only; it gives each leaver its
 *      detection rank, and leavers are staged with their migration shift;
 *   3. per affected cell, arrivals are ordered by
 *      (migration code c2,c1,c0 ; source cell ; detection rank), the scan is
 *      replayed once more to move stayers and early arrivals into place, and
 *      the remaining arrivals are appended.
 * The order key encodes the migration history: an atom that crossed the box
 * in dimension d was appended during phase d of the migration (lower face
 * before upper face), after every atom that reached the same cell earlier.
 *
 * Every kernel body takes (descriptor, logical block).  k_* wrappers run one
 * system; kb_* wrappers run many systems in one launch from a (system,
 * block) work list.
 */
#include "fmdx_internal.h"


#define DEF_KERNELS(NAME)                                                             \
    __global__ __launch_bounds__(TPB) void k_##NAME(KDesc D) { b_##NAME(D, blockIdx.x); } \
    __global__ __launch_bounds__(TPB) void kb_##NAME(BatchMap M)                       \
    {                                                                                  \
        each_item<true>(M, [&](const KDesc &D, int blk) { b_##NAME(D, blk); });         \
    }

static __device__ __forceinline__ void cell_coords(const KDesc &D, int c, int &cx, int &cy, int &cz)
{
    cz = c % D.nc2;
    int t = c / D.nc2;
    cy = t % D.nc1;
    cx = t / D.nc1;
}

/* an integrator's new position: its cell-relative FP32 copy, for the cell the atom is in
   (an atom that leaves gets its copy for the new cell when the re-binning places it) */
static __device__ __forceinline__ void keep_rel(const KDesc &D, int c, long slot, double x, double y, double z)
{
    if (!D.xq_keep) return;
    int cx, cy, cz;
    cell_coords(D, c, cx, cy, cz);
    D.xq[slot] = rel_pos(D, cx, cy, cz, x, y, z);
}

/* destination of one atom after it moved; misc.c:281-300 and md-ghost.c:827-869 */
static __device__ void bin_flag(const KDesc &D, int c, long slot, double x, double y, double z)
{
#pragma clang fp contract(off)
    int cx, cy, cz;
    cell_coords(D, c, cx, cy, cz);
    const double f0 = floor(x / D.h0), f1 = floor(y / D.h1), f2 = floor(z / D.h2);
    /* a non-finite or far-away position cannot be binned (misc.c:281-291); the
       comparisons run in FP64 so that no out-of-range conversion happens */
    if (!(f0 >= -1.0 && f0 <= D.nc0 && f1 >= -1.0 && f1 <= D.nc1 && f2 >= -1.0 && f2 <= D.nc2))
    {
        set_err(D, FMDX_DEV_ERR_POSITION);
        note_first(&D.counters[5], (unsigned)c + 1u);
        D.mv[slot] = MV_STAY;
        return;
    }
    const int j0 = (int)f0 + 1, j1 = (int)f1 + 1, j2 = (int)f2 + 1;
    if (j0 < 0 || j0 >= D.nc0 + 2 || j1 < 0 || j1 >= D.nc1 + 2 || j2 < 0 || j2 >= D.nc2 + 2)
    {
        set_err(D, FMDX_DEV_ERR_POSITION);
        note_first(&D.counters[5], (unsigned)c + 1u);
        D.mv[slot] = MV_STAY;
        return;
    }
    if (j0 == cx + 1 && j1 == cy + 1 && j2 == cz + 1)
    {
        D.mv[slot] = MV_STAY;
        return;
    }
    int f[3], cd[3];
    bool del = false;
    const int j[3] = {j0, j1, j2};
    const int n[3] = {D.nc0, D.nc1, D.nc2};
    const int p[3] = {D.pbc0, D.pbc1, D.pbc2};
    for (int d = 0; d < 3; d++)
    {
        cd[d] = 0;
        if (j[d] == 0)
        {
            if (p[d]) { f[d] = n[d] - 1; cd[d] = 1; }
            else del = true;
        }
        else if (j[d] == n[d] + 1)
        {
            if (p[d]) { f[d] = 0; cd[d] = 2; }
            else del = true;
        }
        else
            f[d] = j[d] - 1;
    }
    int code = cd[2] * 16 + cd[1] * 4 + cd[0];
    atomicAdd(&D.nleave[c], 1);
    D.btouch[c / TPB] = 1;
    D.code[slot] = (uint8_t)code;
    if (del)
    {
        D.mv[slot] = MV_DELETE;
        return;
    }
    int dest = (f[0] * D.nc1 + f[1]) * D.nc2 + f[2];
    D.mv[slot] = dest;
    D.btouch[dest / TPB] = 1;
    atomicAdd(&D.narr[dest], 1);
    if (code == 0 && c < dest) atomicAdd(&D.nearly[dest], 1);
}

/* VelocityVerlet_startStep (integrators.c:40-75) fused with the binning flag: one atom */
static __device__ __forceinline__ void vv_start_atom(const KDesc &D, long slot, int c, double lambda)
{
#pragma clang fp contract(off)
    double x = D.x[slot], y = D.y[slot], z = D.z[slot];
    if (D.ag < 0 || D.grp[slot] == D.ag)
    {
        const double dt = D.dt;
        const double h = D.hdtm[D.kind[slot]];
        double vx = D.vx[slot], vy = D.vy[slot], vz = D.vz[slot];
        if (D.thermostat)
        {
            vx *= lambda;
            vy *= lambda;
            vz *= lambda;
        }
        vx += h * D.fx[slot];
        x += dt * vx;
        vy += h * D.fy[slot];
        y += dt * vy;
        vz += h * D.fz[slot];
        z += dt * vz;
        D.vx[slot] = vx; D.vy[slot] = vy; D.vz[slot] = vz;
        D.x[slot] = x; D.y[slot] = y; D.z[slot] = z;
    }
    keep_rel(D, c, slot, x, y, z);
    bin_flag(D, c, slot, x, y, z);
}

static __device__ __forceinline__ double berendsen_lambda(const KDesc &D)
{
#pragma clang fp contract(off)
    return D.thermostat ? sqrt(1 + D.dt / D.tau * (D.T0 / D.thermo[TH_T] - 1)) : 1.0;
}

static __device__ void b_vv_start(const KDesc &D, int blk)
{
    if (blk == 0 && threadIdx.x == 0) D.counters[0] = 0;
    const double lambda = berendsen_lambda(D);
    each_slot(D, blk, [&](long slot, int c) { vv_start_atom(D, slot, c, lambda); });
}
DEF_KERNELS(vv_start)

/* The start step over a list of occupied (system, cell) items (length *M.count), CPB
   items per workgroup pass, LPC lanes per cell; the workgroups stride over the list.  A
   batch whose members are mostly vacuum has few atoms in each 16-cell group, which the
   group list would visit all of.  The mover counters are reset by kb_occ. */
__global__ __launch_bounds__(TPB) void kb_vv_start_cells(BatchMap M, int *touched)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *touched = 0;   /* kb_touch's list starts empty */
    const int n = *M.count;
    const int sub = (int)threadIdx.x / LPC, l = (int)threadIdx.x % LPC;
    for (int i = (int)blockIdx.x * CPB + sub; i - sub < n; i += (int)gridDim.x * CPB)
    {
        if (i >= n) continue;
        const int2 it = M.items[i];
        const KDesc &D = M.descs[it.x];
        const int c = it.y;
        const double lambda = berendsen_lambda(D);
        const int cnt = D.cnt[c];
        for (int k = l; k < cnt; k += LPC) vv_start_atom(D, (long)c * D.cap + k, c, lambda);
    }
}

/* SymplecticEuler_takeOneStep, first half (integrators.c:134-147) + binning flag */
static __device__ void b_se_step(const KDesc &D, int blk)
{
#pragma clang fp contract(off)
    if (blk == 0 && threadIdx.x == 0) D.counters[0] = 0;
    each_slot(D, blk, [&](long slot, int c) {
        double x = D.x[slot], y = D.y[slot], z = D.z[slot];
        if (D.ag < 0 || D.grp[slot] == D.ag)
        {
            const double dt = D.dt;
            const double m = D.mass[D.kind[slot]];
            double vx = D.vx[slot], vy = D.vy[slot], vz = D.vz[slot];
            vx += dt * D.fx[slot] / m;
            x += dt * vx;
            vy += dt * D.fy[slot] / m;
            y += dt * vy;
            vz += dt * D.fz[slot] / m;
            z += dt * vz;
            D.vx[slot] = vx; D.vy[slot] = vy; D.vz[slot] = vz;
            D.x[slot] = x; D.y[slot] = y; D.z[slot] = z;
        }
        keep_rel(D, c, slot, x, y, z);
        bin_flag(D, c, slot, x, y, z);
    });
}
DEF_KERNELS(se_step)

/* binning flag without motion (_fmd_refreshGrid called on its own) */
static __device__ void b_flag_only(const KDesc &D, int blk)
{
    if (blk == 0 && threadIdx.x == 0) D.counters[0] = 0;
    each_slot(D, blk, [&](long slot, int c) { bin_flag(D, c, slot, D.x[slot], D.y[slot], D.z[slot]); });
}
DEF_KERNELS(flag_only)

/* per-block deterministic tree sum of 5 values into part[blk*8 + q] */
static __device__ void block_partials(const KDesc &D, int blk, double (&q)[5])
{
    __shared__ double sh[5][TPB];
    for (int i = 0; i < 5; i++) sh[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            for (int i = 0; i < 5; i++) sh[i][threadIdx.x] += sh[i][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int i = 0; i < 5; i++) D.part[(long)blk * 8 + i] = sh[i][0];
}

/* VelocityVerlet_finishStep, integrators.c:94-126: second kick + group sums */
static __device__ void b_vv_finish(const KDesc &D, int blk)
{
#pragma clang fp contract(off)
    double q[5] = {0, 0, 0, 0, 0};
    each_slot(D, blk, [&](long slot, int c) {
        (void)c;
        if (D.ag < 0 || D.grp[slot] == D.ag)
        {
            const int kd = D.kind[slot];
            const double h = D.hdtm[kd], m = D.mass[kd];
            const double vx = D.vx[slot] + h * D.fx[slot];
            const double vy = D.vy[slot] + h * D.fy[slot];
            const double vz = D.vz[slot] + h * D.fz[slot];
            D.vx[slot] = vx; D.vy[slot] = vy; D.vz[slot] = vz;
            q[0] += m * (vx * vx + vy * vy + vz * vz);
            q[1] += m * vx;
            q[2] += m * vy;
            q[3] += m * vz;
            q[4] += 1.0;
        }
    });
    block_partials(D, blk, q);
}
DEF_KERNELS(vv_finish)

/* group sums without a kick: matter.c:31-68 and integrators.c:154-176 */
static __device__ void b_thermo(const KDesc &D, int blk)
{
#pragma clang fp contract(off)
    double q[5] = {0, 0, 0, 0, 0};
    each_slot(D, blk, [&](long slot, int c) {
        (void)c;
        if (D.ag < 0 || D.grp[slot] == D.ag)
        {
            const double m = D.mass[D.kind[slot]];
            const double vx = D.vx[slot], vy = D.vy[slot], vz = D.vz[slot];
            q[0] += m * (vx * vx + vy * vy + vz * vz);
            q[1] += m * vx;
            q[2] += m * vy;
            q[3] += m * vz;
            q[4] += 1.0;
        }
    });
    block_partials(D, blk, q);
}
DEF_KERNELS(thermo)

static __device__ __forceinline__ int nblk(const KDesc &D) { return (D.ncell + CPB - 1) / CPB; }

/* fixed-order reduction of the per-block partials: updateGroupProperties, integrators.c:78-92 */
static __device__ void b_reduce_thermo(const KDesc &D)
{
#pragma clang fp contract(off)
    __shared__ double sh[5][1024];
    const int nb = nblk(D);
    double q[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 1024)
        for (int i = 0; i < 5; i++) q[i] += D.part[(long)b * 8 + i];
    for (int i = 0; i < 5; i++) sh[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            for (int i = 0; i < 5; i++) sh[i][threadIdx.x] += sh[i][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        double mv2 = sh[0][0];
        double n = sh[4][0];
        D.thermo[TH_MV2] = mv2;
        D.thermo[TH_KE] = 0.5 * mv2;
        D.thermo[TH_T] = mv2 / (3.0 * n * KB_EV);
        D.thermo[TH_PX] = sh[1][0];
        D.thermo[TH_PY] = sh[2][0];
        D.thermo[TH_PZ] = sh[3][0];
        D.thermo[TH_N] = n;
    }
}
__global__ __launch_bounds__(1024) void k_reduce_thermo(KDesc D) { b_reduce_thermo(D); }
__global__ __launch_bounds__(1024) void kb_reduce_thermo(const KDesc *descs)
{
    const KDesc &D = descs[blockIdx.x];
    b_reduce_thermo(D);
}


/* two-level fixed-order sum of the per-cell energies of a system: block b of
   nb sums a contiguous run of cells, one block sums the nb block totals */
static __device__ void b_pe_part(const KDesc &D, int blk, int nb)
{
    __shared__ double sh[2][TPB];
    const int per = (D.ncell + nb - 1) / nb;
    const int c0 = blk * per, c1 = min(D.ncell, c0 + per);
    double pe = 0, emb = 0;
    for (int c = c0 + threadIdx.x; c < c1; c += TPB)
    {
        pe += D.cell_pe[c];
        emb += D.cell_emb[c];
    }
    sh[0][threadIdx.x] = pe;
    sh[1][threadIdx.x] = emb;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
        {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        D.epart[2 * blk] = sh[0][0];
        D.epart[2 * blk + 1] = sh[1][0];
    }
}
__global__ __launch_bounds__(TPB) void k_pe_part(KDesc D) { b_pe_part(D, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(TPB) void kb_pe_part(BatchMap M)
{
    const int2 it = M.items[blockIdx.x];
    const KDesc &D = M.descs[it.x];
    b_pe_part(D, it.y, pe_blocks(D.ncell));
}

static __device__ void b_pe_final(const KDesc &D, int nb)
{
    __shared__ double sh[2][1024];
    double pe = 0, emb = 0;
    for (int b = threadIdx.x; b < nb; b += 1024)
    {
        pe += D.epart[2 * b];
        emb += D.epart[2 * b + 1];
    }
    sh[0][threadIdx.x] = pe;
    sh[1][threadIdx.x] = emb;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
        {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        D.thermo[TH_PE] = sh[0][0] + sh[1][0];
        D.thermo[TH_EMB] = sh[1][0];
    }
}
__global__ __launch_bounds__(1024) void k_pe_final(KDesc D, int nb) { b_pe_final(D, nb); }
__global__ __launch_bounds__(1024) void kb_pe_final(const KDesc *descs)
{
    const KDesc &D = descs[blockIdx.x];
    b_pe_final(D, pe_blocks(D.ncell));
}

/* status words and thermo records of every system of a batch, for one copy to the host */
__global__ __launch_bounds__(64) void kb_collect(const KDesc *descs, unsigned *status, double *thermo)
{
    const KDesc &D = descs[blockIdx.x];
    if (threadIdx.x < 8) status[blockIdx.x * 8 + threadIdx.x] = D.counters[threadIdx.x];
    if (threadIdx.x < 16) thermo[blockIdx.x * 16 + threadIdx.x] = D.thermo[threadIdx.x];
}

/* ----------------------------------------------------------- binning */

/* a stayer moved within its cell: its cell-relative copy moves with it */
static __device__ __forceinline__ void copy_slot(const KDesc &D, long from, long to)
{
    D.x[to] = D.x[from]; D.y[to] = D.y[from]; D.z[to] = D.z[from];
    D.vx[to] = D.vx[from]; D.vy[to] = D.vy[from]; D.vz[to] = D.vz[from];
    D.grp[to] = D.grp[from]; D.aid[to] = D.aid[from]; D.kind[to] = D.kind[from];
    if (D.xq_keep) D.xq[to] = D.xq[from];
}

/* an arrival (or a stayer staged with the early arrivals) placed in cell (cx, cy, cz):
   its cell-relative copy is formed for that cell */
static __device__ __forceinline__ void copy_staged(const KDesc &D, int mid, long to, int cx, int cy, int cz)
{
    D.x[to] = D.mvx[mid]; D.y[to] = D.mvy[mid]; D.z[to] = D.mvz[mid];
    D.vx[to] = D.mvvx[mid]; D.vy[to] = D.mvvy[mid]; D.vz[to] = D.mvvz[mid];
    D.grp[to] = D.mvgrp[mid]; D.aid[to] = D.mvaid[mid]; D.kind[to] = D.mvkind[mid];
    if (D.xq_keep) D.xq[to] = rel_pos(D, cx, cy, cz, D.mvx[mid], D.mvy[mid], D.mvz[mid]);
}

/* The replays below walk a cell's move flags in a data-dependent order.  A cell of up to
   64 atoms first turns its flags into a bit mask (bit k: slot k leaves) with independent
   16-byte loads, two batches of up to 8 (slot capacities that are multiples of 4 keep the
   loads aligned and inside the cell); larger cells, and capacities set by a caller to other
   values, read the flags from HBM at each step of the walk.  The walks only decide which
   atom goes where: the atoms are then staged or moved by all threads of the block, one
   atom each, so that no thread carries a chain of dependent copies (a slowest cell's chain
   set the time of the whole re-binning). */
struct MoveFlags
{
    const int *g;
    unsigned long long m;
    bool local;
    __device__ __forceinline__ bool leaves(int k) const
    {
        return local ? ((m >> k) & 1ull) != 0ull : g[k] != MV_STAY;
    }
};

static __device__ __forceinline__ MoveFlags move_flags(const KDesc &D, long base, int nc)
{
    MoveFlags f;
    f.g = D.mv + base;
    f.m = 0ull;
    f.local = nc <= 64 && (D.cap & 3) == 0;
    if (f.local)
    {
        const int4 *src = reinterpret_cast<const int4 *>(D.mv + base);
        const int nq = (nc + 3) >> 2;   /* 4 nq <= cap: the capacity is a multiple of 4 and holds nc */
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            int4 q[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (8 * h + k < nq) q[k] = src[8 * h + k];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (8 * h + k < nq)
                    f.m |= (unsigned long long)((unsigned)(q[k].x != MV_STAY) | ((unsigned)(q[k].y != MV_STAY) << 1) |
                                                ((unsigned)(q[k].z != MV_STAY) << 2) |
                                                ((unsigned)(q[k].w != MV_STAY) << 3))
                           << (4 * (8 * h + k));
        }
    }
    return f;
}

#define WAVE 64   /* lanes per wavefront on gfx950 */

/* exclusive prefix of v over the block, and the block total, through the waves' shuffles */
static __device__ __forceinline__ unsigned block_prefix(unsigned v, unsigned &total)
{
    __shared__ unsigned wsum[TPB / WAVE];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    unsigned incl = v;
    for (int o = 1; o < WAVE; o <<= 1)
    {
        const unsigned t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    unsigned before = 0;
    total = 0;
    for (int k = 0; k < TPB / WAVE; k++)
    {
        if (k < w) before += wsum[k];
        total += wsum[k];
    }
    __syncthreads();
    return before + incl - v;
}

/* one leaver of cell c at slot s into staging entry mid, rank-th leaver of its cell: its
   position as ccopy_for_migrate leaves it and its arrival key in the destination's bucket */
static __device__ __forceinline__ void stage_leaver(const KDesc &D, int c, long s, int mid, int rank)
{
#pragma clang fp contract(off)
    const int dst = D.mv[s];
    if (dst == MV_DELETE)
    {
        atomicAdd(&D.counters[1], 1u);
        return;
    }
    const int code = D.code[s];
    double x = D.x[s], y = D.y[s], z = D.z[s];
    /* ccopy_for_migrate, md-ghost.c:132-140: x += L (lower face) or x += -L (upper) */
    if ((code & 3) == 1) x += D.L0;
    else if ((code & 3) == 2) x += -D.L0;
    if (((code >> 2) & 3) == 1) y += D.L1;
    else if (((code >> 2) & 3) == 2) y += -D.L1;
    if (((code >> 4) & 3) == 1) z += D.L2;
    else if (((code >> 4) & 3) == 2) z += -D.L2;
    D.mvx[mid] = x; D.mvy[mid] = y; D.mvz[mid] = z;
    D.mvvx[mid] = D.vx[s]; D.mvvy[mid] = D.vy[s]; D.mvvz[mid] = D.vz[s];
    D.mvgrp[mid] = D.grp[s]; D.mvaid[mid] = D.aid[s]; D.mvkind[mid] = D.kind[s];
    const unsigned long long key = ((unsigned long long)code << 58) | ((unsigned long long)c << 16) |
                                   (unsigned long long)rank;
    const int pos = atomicAdd(&D.fill[dst], 1);
    if (pos < D.cap)
    {
        D.bkey[(long)dst * D.cap + pos] = key;
        D.bid[(long)dst * D.cap + pos] = mid;
    }
    else
        set_err(D, FMDX_DEV_ERR_OVERFLOW);
}

/* replay of the swap-remove scan (misc.c:279-303) to rank and stage leavers.
   Each block reserves one run of staging entries for all its cells with a
   single atomic (a cell takes one entry per leaver, deletions included, so
   the staging buffer of natoms entries always suffices); one atomic per
   leaver on the shared counter serialised across the whole device.  The walks
   list the block's leavers in staging order; all threads then stage them (a
   block with more than LVMAX leavers stages them in its walks). */
#define LVMAX 1024
static __device__ void b_bin_rank(const KDesc &D, int blk)
{
    __shared__ int lv_slot[LVMAX];        /* slot of the leaver within its cell */
    __shared__ uint8_t lv_cell[LVMAX];    /* its cell within the block */
    __shared__ unsigned lv_off[TPB];      /* each cell's first entry of the block's run */
    __shared__ unsigned run0;
    const int c = blk * TPB + threadIdx.x;
    const int nl = c < D.ncell ? D.nleave[c] : 0;
    unsigned total;
    const unsigned off = block_prefix((unsigned)nl, total);
    lv_off[threadIdx.x] = off;
    if (threadIdx.x == 0) run0 = total ? atomicAdd(&D.counters[0], total) : 0u;
    __syncthreads();
    if (total == 0) return;
    const bool listed = total <= LVMAX;
    if (nl != 0)
    {
        const int nc = D.cnt[c];
        const long base = (long)c * D.cap;
        const MoveFlags mv = move_flags(D, base, nc);
        int i = 0, n = nc + D.nearly[c], cur = 0, rank = 0;
        while (i < n)
        {
            if (cur < nc && mv.leaves(cur))
            {
                if (listed)
                {
                    lv_slot[off + rank] = cur;
                    lv_cell[off + rank] = (uint8_t)threadIdx.x;
                }
                else
                    stage_leaver(D, c, base + cur, (int)(run0 + off + rank), rank);
                rank++;
                n--;
                if (n != i) cur = n;
            }
            else
            {
                i++;
                cur = i;
            }
        }
    }
    if (!listed) return;
    __syncthreads();
    for (unsigned j = threadIdx.x; j < total; j += TPB)
    {
        const int t = lv_cell[j];
        const int cc = blk * TPB + t;
        stage_leaver(D, cc, (long)cc * D.cap + lv_slot[j], (int)(run0 + j), (int)(j - lv_off[t]));
    }
}
DEF_KERNELS(bin_rank)

/* place stayers, early and late arrivals of every affected cell: the walks sort each
   cell's arrival bucket (in LDS when the block's arrivals fit ARMAX) and list the cell's
   moves, which all threads then carry out in two stages: first the moves within the
   cell's surviving prefix (stayers into holes, early arrivals), whose sources (slots at or
   beyond the prefix, or staging) and destinations never meet, then the late arrivals
   appended after the prefix, which may land on slots the first stage read.  A cell whose
   moves do not fit the block's list (MOVMAX) makes them itself, in the walk's order. */
#define ARMAX 1024
#define MOVMAX 2048
struct MoveList
{
    int2 *m;            /* (cell within block | destination slot << 8 | staged << 20 | late << 21, source),
                           or x = -1: an entry of a reservation that did not fit */
    unsigned *n;
};

static __device__ int bin_apply_cell(const KDesc &D, int c, int nl, int na, unsigned long long *akey, int *amid,
                                     const MoveList &L);

static __device__ __forceinline__ void do_move(const KDesc &D, int blk, int2 m)
{
    const int c = blk * TPB + (m.x & 255);
    const long base = (long)c * D.cap;
    const int dst = (m.x >> 8) & 0xfff;
    if ((m.x >> 20) & 1)
    {
        int cx, cy, cz;
        cell_coords(D, c, cx, cy, cz);
        copy_staged(D, m.y, base + dst, cx, cy, cz);
    }
    else
        copy_slot(D, base + m.y, base + dst);
}

static __device__ void b_bin_apply(const KDesc &D, int blk)
{
    __shared__ unsigned long long akey[ARMAX];
    __shared__ int amid[ARMAX];
    __shared__ int2 mov[MOVMAX];
    __shared__ unsigned nmov;
    const int c = blk * TPB + threadIdx.x;
    const int nl = c < D.ncell ? D.nleave[c] : 0, na = c < D.ncell ? D.narr[c] : 0;
    const int nb = na < D.cap ? na : D.cap;
    unsigned atot;
    const unsigned aoff = block_prefix((unsigned)nb, atot);
    if (threadIdx.x == 0) nmov = 0u;
    __syncthreads();
    const bool lds_sort = atot <= ARMAX;
    const MoveList L = {mov, &nmov};
    int total = 0;
    if (nl != 0 || na != 0)
        total = bin_apply_cell(D, c, nl, na, lds_sort ? akey + aoff : nullptr, lds_sort ? amid + aoff : nullptr, L);
    __syncthreads();
    const unsigned nm = nmov < MOVMAX ? nmov : MOVMAX;
    for (int late = 0; late < 2; late++)
    {
        if (late) __syncthreads();   /* the slots the late arrivals land on have been read */
        for (unsigned j = threadIdx.x; j < nm; j += TPB)
        {
            const int2 m = mov[j];
            if (m.x >= 0 && ((m.x >> 21) & 1) == late) do_move(D, blk, m);
        }
    }
    /* the fullest cell: one atomic per block (a per-cell atomic on the shared word serialises) */
    unsigned m = (unsigned)total;
    for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor(m, o));
    __shared__ unsigned wmax[TPB / WAVE];
    if (threadIdx.x % WAVE == 0) wmax[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned bm = 0;
        for (int k = 0; k < TPB / WAVE; k++) bm = max(bm, wmax[k]);
        if (bm) atomicMax(&D.counters[2], bm);
    }
}

static __device__ int bin_apply_cell(const KDesc &D, int c, int nl, int na, unsigned long long *akey, int *amid,
                                     const MoveList &L)
{
    int cx, cy, cz;
    cell_coords(D, c, cx, cy, cz);
    const int nc = D.cnt[c];
    const int A = D.nearly[c];
    const long base = (long)c * D.cap;
    const MoveFlags mv = move_flags(D, base, nc);
    const int nb = na < D.cap ? na : D.cap;
    /* insertion sort of the arrival bucket by key */
    if (akey)
    {
        for (int a = 0; a < nb; a++)
        {
            akey[a] = D.bkey[base + a];
            amid[a] = D.bid[base + a];
        }
        for (int a = 1; a < nb; a++)
        {
            const unsigned long long k = akey[a];
            const int id = amid[a];
            int b = a - 1;
            while (b >= 0 && akey[b] > k)
            {
                akey[b + 1] = akey[b];
                amid[b + 1] = amid[b];
                b--;
            }
            akey[b + 1] = k;
            amid[b + 1] = id;
        }
    }
    else
        for (int a = 1; a < nb; a++)
        {
            unsigned long long k = D.bkey[base + a];
            int id = D.bid[base + a];
            int b = a - 1;
            while (b >= 0 && D.bkey[base + b] > k)
            {
                D.bkey[base + b + 1] = D.bkey[base + b];
                D.bid[base + b + 1] = D.bid[base + b];
                b--;
            }
            D.bkey[base + b + 1] = k;
            D.bid[base + b + 1] = id;
        }
    auto staged = [&](int q) { return akey ? amid[q] : D.bid[base + q]; };
    /* the replay (misc.c:279-303 with the early arrivals appended): act(destination, staged,
       source, late) for every atom that changes slot; returns the count before the late
       arrivals */
    auto walk = [&](auto act) {
        int i = 0, n = nc + A, cur = 0;
        while (i < n)
        {
            if (cur < nc && mv.leaves(cur))
            {
                n--;
                if (n != i) cur = n;
            }
            else
            {
                if (cur >= nc)
                {
                    int q = cur - nc;
                    if (q < nb && i < D.cap) act(i, true, q, false);
                }
                else if (cur != i)
                    act(i, false, cur, false);
                i++;
                cur = i;
            }
        }
        int t = n;
        for (int q = A; q < na; q++)
        {
            if (q < nb && t < D.cap) act(t, true, q, true);
            t++;
        }
        return t;
    };
    int nmv = 0;
    const int total0 = walk([&](int, bool, int, bool) { nmv++; });
    if (nmv)
    {
        const unsigned j0 = atomicAdd(L.n, (unsigned)nmv);
        if (j0 + (unsigned)nmv <= MOVMAX)
        {
            unsigned j = j0;
            walk([&](int dst, bool st, int src, bool late) {
                L.m[j++] = make_int2((int)(threadIdx.x & 255) | (dst << 8) | (st ? 1 << 20 : 0) | (late ? 1 << 21 : 0),
                                     st ? staged(src) : src);
            });
        }
        else
        {
            for (unsigned j = j0; j < MOVMAX; j++) L.m[j] = make_int2(-1, 0);
            walk([&](int dst, bool st, int src, bool) {
                if (st) copy_staged(D, staged(src), base + dst, cx, cy, cz);
                else copy_slot(D, base + src, base + dst);
            });
        }
    }
    int total = total0;
    if (total > D.cap)
    {
        set_err(D, FMDX_DEV_ERR_OVERFLOW);
        total = D.cap;
    }
    D.cnt[c] = total;
    D.nleave[c] = 0;
    D.narr[c] = 0;
    D.nearly[c] = 0;
    D.fill[c] = 0;
    return total;
}
DEF_KERNELS(bin_apply)

__global__ __launch_bounds__(TPB) void k_add_velocity(KDesc D, int group, double vx, double vy, double vz)
{
#pragma clang fp contract(off)
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        (void)c;
        if (group == -1 || D.grp[slot] == group)
        {
            D.vx[slot] += vx;
            D.vy[slot] += vy;
            D.vz[slot] += vz;
        }
    });
}

/* copy every cell into a layout with another slot capacity */
__global__ __launch_bounds__(TPB) void k_relayout(KDesc A, KDesc B)
{
    each_slot(A, blockIdx.x, [&](long slot, int c) {
        const int k = (int)(slot - (long)c * A.cap);
        if (k >= B.cap) return;
        const long t = (long)c * B.cap + k;
        B.x[t] = A.x[slot]; B.y[t] = A.y[slot]; B.z[t] = A.z[slot];
        B.vx[t] = A.vx[slot]; B.vy[t] = A.vy[slot]; B.vz[t] = A.vz[slot];
        B.fx[t] = A.fx[slot]; B.fy[t] = A.fy[slot]; B.fz[t] = A.fz[slot];
        B.aux[t] = A.aux[slot];
        B.xq[t] = A.xq[slot];
        B.grp[t] = A.grp[slot]; B.aid[t] = A.aid[slot]; B.kind[t] = A.kind[slot];
    });
}

/* compact cell-ordered arrays -> slots (transfer_from_globalgrid_to_rank0_grid, matter.c:329-363) */
__global__ __launch_bounds__(TPB) void k_scatter(KDesc D, const int *cell_start, const double *x, const double *v,
                                                  const int *grp, const unsigned *aid, const uint8_t *kind)
{
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        const long a = (long)cell_start[c] + (slot - (long)c * D.cap);
        D.x[slot] = x[3 * a]; D.y[slot] = x[3 * a + 1]; D.z[slot] = x[3 * a + 2];
        D.vx[slot] = v[3 * a]; D.vy[slot] = v[3 * a + 1]; D.vz[slot] = v[3 * a + 2];
        D.fx[slot] = 0.0; D.fy[slot] = 0.0; D.fz[slot] = 0.0; D.aux[slot] = 0.0;
        D.grp[slot] = grp[a]; D.aid[slot] = aid[a]; D.kind[slot] = kind[a];
    });
}

/* slots -> compact cell-ordered arrays */
__global__ __launch_bounds__(TPB) void k_gather(KDesc D, const int *cell_start, double *x, double *v, double *f,
                                                 double *aux, int *grp, unsigned *aid, uint8_t *kind)
{
    each_slot(D, blockIdx.x, [&](long slot, int c) {
        const long a = (long)cell_start[c] + (slot - (long)c * D.cap);
        x[3 * a] = D.x[slot]; x[3 * a + 1] = D.y[slot]; x[3 * a + 2] = D.z[slot];
        v[3 * a] = D.vx[slot]; v[3 * a + 1] = D.vy[slot]; v[3 * a + 2] = D.vz[slot];
        f[3 * a] = D.fx[slot]; f[3 * a + 1] = D.fy[slot]; f[3 * a + 2] = D.fz[slot];
        aux[a] = D.aux[slot];
        grp[a] = D.grp[slot]; aid[a] = D.aid[slot]; kind[a] = D.kind[slot];
    });
}

/* state check between steps: atoms over the cells (counters[8]), the fullest cell ([9]) and the
   atoms whose position, velocity or force is not finite ([10]); a grid-stride walk over groups of
   CPB cells, one set of atomics per block */
__global__ __launch_bounds__(TPB) void k_audit(KDesc D)
{
    unsigned n = 0, m = 0, bad = 0;
    const int ng = (D.ncell + CPB - 1) / CPB;
    for (int g = blockIdx.x; g < ng; g += gridDim.x)
    {
        const int c = g * CPB + (int)threadIdx.x / LPC;
        if (c >= D.ncell) continue;
        const int k0 = (int)threadIdx.x % LPC;
        const int cnt = D.cnt[c];
        if (k0 == 0)
        {
            n += (unsigned)cnt;
            m = max(m, (unsigned)cnt);
        }
        const int lim = min(cnt, D.cap);
        for (int k = k0; k < lim; k += LPC)
        {
            const long q = (long)c * D.cap + k;
            /* the sum is finite only if every term is (or overflows, which is as bad) */
            const double t = D.x[q] + D.y[q] + D.z[q] + D.vx[q] + D.vy[q] + D.vz[q] + D.fx[q] + D.fy[q] + D.fz[q];
            if (!__builtin_isfinite(t)) bad++;
        }
    }
    __shared__ unsigned sh[3][TPB];
    sh[0][threadIdx.x] = n;
    sh[1][threadIdx.x] = m;
    sh[2][threadIdx.x] = bad;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1)
    {
        if ((int)threadIdx.x < s)
        {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] = max(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        atomicAdd(&D.counters[8], sh[0][0]);
        atomicMax(&D.counters[9], sh[1][0]);
        atomicAdd(&D.counters[10], sh[2][0]);
    }
}

__global__ __launch_bounds__(64) void k_audit_clear(KDesc D)
{
    if (threadIdx.x >= 8 && threadIdx.x <= 10) D.counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(TPB) void k_fill(double *p, long n, double v)
{
    long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

/* ------------------------------------------------------------ launchers */

static inline unsigned nblk_h(const KDesc &D) { return (unsigned)((D.ncell + CPB - 1) / CPB); }
static inline unsigned ncblk_h(const KDesc &D) { return (unsigned)((D.ncell + TPB - 1) / TPB); }

hipError_t launch_bin(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_bin_rank, dim3(ncblk_h(D)), dim3(TPB), 0, st, D);
    hipLaunchKernelGGL(k_bin_apply, dim3(ncblk_h(D)), dim3(TPB), 0, st, D);
    return hipGetLastError();
}

hipError_t launch_vv_start(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_vv_start, dim3(nblk_h(D)), dim3(TPB), 0, st, D);
    return launch_bin(D, st);
}

hipError_t launch_se_step(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_se_step, dim3(nblk_h(D)), dim3(TPB), 0, st, D);
    hipError_t e = launch_bin(D, st);
    if (e != hipSuccess) return e;
    return launch_thermo(D, st);
}

hipError_t launch_flag_only(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_flag_only, dim3(nblk_h(D)), dim3(TPB), 0, st, D);
    return launch_bin(D, st);
}

hipError_t launch_vv_finish(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_vv_finish, dim3(nblk_h(D)), dim3(TPB), 0, st, D);
    hipLaunchKernelGGL(k_reduce_thermo, dim3(1), dim3(1024), 0, st, D);
    return hipGetLastError();
}

hipError_t launch_thermo(const KDesc &D, hipStream_t st)
{
    hipLaunchKernelGGL(k_thermo, dim3(nblk_h(D)), dim3(TPB), 0, st, D);
    hipLaunchKernelGGL(k_reduce_thermo, dim3(1), dim3(1024), 0, st, D);
    return hipGetLastError();
}

hipError_t launch_reduce_pe(const KDesc &D, hipStream_t st)
{
    const int nb = pe_blocks(D.ncell);
    hipLaunchKernelGGL(k_pe_part, dim3(nb), dim3(TPB), 0, st, D);
    hipLaunchKernelGGL(k_pe_final, dim3(1), dim3(1024), 0, st, D, nb);
    return hipGetLastError();
}

hipError_t launch_add_velocity(const KDesc &D, hipStream_t st, int group, double vx, double vy, double vz)
{
    hipLaunchKernelGGL(k_add_velocity, dim3(nblk_h(D)), dim3(TPB), 0, st, D, group, vx, vy, vz);
    return hipGetLastError();
}

hipError_t launch_relayout(const KDesc &A, const KDesc &B, hipStream_t st)
{
    hipLaunchKernelGGL(k_relayout, dim3(nblk_h(A)), dim3(TPB), 0, st, A, B);
    return hipGetLastError();
}

hipError_t launch_scatter(const KDesc &D, hipStream_t st, const int *cell_start, const double *x, const double *v,
                          const int *grp, const unsigned *aid, const uint8_t *kind)
{
    hipLaunchKernelGGL(k_scatter, dim3(nblk_h(D)), dim3(TPB), 0, st, D, cell_start, x, v, grp, aid, kind);
    return hipGetLastError();
}

hipError_t launch_gather(const KDesc &D, hipStream_t st, const int *cell_start, double *x, double *v, double *f,
                         double *aux, int *grp, unsigned *aid, uint8_t *kind)
{
    hipLaunchKernelGGL(k_gather, dim3(nblk_h(D)), dim3(TPB), 0, st, D, cell_start, x, v, f, aux, grp, aid, kind);
    return hipGetLastError();
}

hipError_t launch_audit(const KDesc &D, hipStream_t st)
{
    const unsigned ng = nblk_h(D);
    hipLaunchKernelGGL(k_audit_clear, dim3(1), dim3(64), 0, st, D);
    hipLaunchKernelGGL(k_audit, dim3(ng < 1024 ? (ng ? ng : 1) : 1024), dim3(TPB), 0, st, D);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, long n, double v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, p, n, v);
    return hipGetLastError();
}

/* ------------------------------------------- occupied cells of a batch */

/* zero the list lengths whose bits are set; a kernel, so that a captured step holds no memset */
__global__ __launch_bounds__(64) void kb_occ_clear(int *occ_n, unsigned mask)
{
    if (threadIdx.x < 16 && ((mask >> threadIdx.x) & 1u)) occ_n[threadIdx.x] = 0;
}

/* Before a re-binning: the (system, block of TPB cells) items whose cells lose or gain
   atoms, as the start step flagged them, into occ_bin (length occ_n[9], zeroed by the start
   step; it stays until the next step, for the status read); the flags are cleared.  The re-binning kernels then skip the blocks
   where nothing moves, which they would leave as they are. */
__global__ __launch_bounds__(TPB) void kb_touch(const KDesc *descs, const int2 *items, int n, int2 *list, int *occ_n)
{
    int *const len = occ_n + 9;
    /* the lists kb_occ is about to write start empty (their readers of the last step and
       the start step, which read occ_n[10], have finished) */
    if (blockIdx.x == 0 && threadIdx.x < 11 && threadIdx.x != 9) occ_n[threadIdx.x] = 0;
    const int i = (int)(blockIdx.x * TPB + threadIdx.x);
    bool t = false;
    int2 it = make_int2(0, 0);
    if (i < n)
    {
        it = items[i];
        int *const f = &descs[it.x].btouch[it.y];
        t = *f != 0;
        if (t) *f = 0;
    }
    const unsigned long long m = __ballot(t);
    if (!m) return;
    const int lane = (int)(threadIdx.x % WAVE);
    const int first = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(len, __popcll(m));
    base = __shfl(base, first);
    if (t) list[base + __popcll(m & ((1ull << lane) - 1ull))] = it;
}

/* After a re-binning, over the (system, block of TPB cells) items, OCC_K consecutive
   items per workgroup: each occupied cell into its system's force-pass list (and
   density-pass list, EAM) and the list of all occupied cells (the start step), each group
   of CPB cells holding an atom into the O(N) list.
   An empty cell gets the words the force passes write for one (no pair or embedding
   energy, no recorded hits) and an empty group the zero thermo partials the kick writes
   for it, so that the reductions over every cell and group read what they read without
   the compaction; each system's mover counter is reset for the next re-binning.  A
   workgroup reserves its run of each list it adds to with one atomic (the reservations
   serialise on the list lengths) and writes its items in cell order.  The order of a
   list decides only which workgroup takes an item, never a result. */
#define OCC_K 4
__global__ __launch_bounds__(TPB) void kb_occ(BatchMap M, BatchPlan P)
{
    __shared__ int wc[OCC_K][2][TPB / WAVE];
    __shared__ int run[OCC_K][4];
    __shared__ int lst[OCC_K][4];   /* list index of each item's four appends, -1: none */
    const int lane = (int)(threadIdx.x % WAVE), w = (int)(threadIdx.x / WAVE);
    const bool lead = (lane % CPB) == 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    int n[OCC_K];
    unsigned long long occ[OCC_K], gm[OCC_K];
    int2 it[OCC_K];
#pragma unroll
    for (int k = 0; k < OCC_K; k++)
    {
        const int idx = (int)blockIdx.x * OCC_K + k;
        n[k] = 0;
        occ[k] = gm[k] = 0ull;
        it[k] = make_int2(-1, 0);
        if (idx < P.n_cellblk)   /* uniform */
        {
            it[k] = M.items[idx];
            const KDesc &D = P.descs[it[k].x];
            const int c = it[k].y * TPB + (int)threadIdx.x;
            const bool in = c < D.ncell;
            n[k] = in ? D.cnt[c] : 0;
            if (in && n[k] == 0)
            {
                D.cell_pe[c] = 0.0;
                if (D.has_eam)
                {
                    D.cell_emb[c] = 0.0;
                    D.htrips[c] = -1;
                }
            }
            if (it[k].y == 0 && threadIdx.x == 0) D.counters[0] = 0;
            /* the group of lane l is lanes l & ~(CPB-1) .. | (CPB-1) of the same wave; a group
               exists when its first cell does */
            occ[k] = __ballot(n[k] > 0);
            const bool gocc = ((occ[k] >> (lane & ~(CPB - 1))) & ((1ull << CPB) - 1ull)) != 0ull;
            gm[k] = __ballot(lead && in && gocc);
            if (lead && in && !gocc)
                for (int i = 0; i < 5; i++) D.part[(long)(c / CPB) * 8 + i] = 0.0;
            if (threadIdx.x == 0)
            {
                const int cat = P.sys_cat[it[k].x];
                lst[k][0] = 1 + cat;
                lst[k][1] = D.has_eam ? 5 + cat : -1;
                lst[k][2] = 0;
                lst[k][3] = 10;
            }
        }
        else if (threadIdx.x == 0)
            lst[k][0] = lst[k][1] = lst[k][2] = lst[k][3] = -1;
        if (lane == 0)
        {
            wc[k][0][w] = __popcll(occ[k]);
            wc[k][1][w] = __popcll(gm[k]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        /* each item's count per append, then one reservation per list the workgroup adds to */
        int cnt[OCC_K][4];
        for (int k = 0; k < OCC_K; k++)
        {
            int t0 = 0, t1 = 0;
            for (int v = 0; v < TPB / WAVE; v++)
            {
                t0 += wc[k][0][v];
                t1 += wc[k][1][v];
            }
            cnt[k][0] = t0;
            cnt[k][1] = t0;
            cnt[k][2] = t1;
            cnt[k][3] = t0;
        }
        for (int k = 0; k < OCC_K; k++)
            for (int a = 0; a < 4; a++)
            {
                const int l = lst[k][a];
                if (l < 0) continue;
                /* first item with this list: reserve for every item that shares it */
                bool seen = false;
                for (int k2 = 0; k2 < k; k2++)
                    for (int a2 = 0; a2 < 4; a2++) seen |= lst[k2][a2] == l;
                if (seen) continue;
                int tot = 0;
                for (int k2 = k; k2 < OCC_K; k2++)
                    for (int a2 = 0; a2 < 4; a2++)
                        if (lst[k2][a2] == l) tot += cnt[k2][a2];
                int at = tot ? atomicAdd(&P.occ_n[l], tot) : 0;
                for (int k2 = k; k2 < OCC_K; k2++)
                    for (int a2 = 0; a2 < 4; a2++)
                        if (lst[k2][a2] == l)
                        {
                            run[k2][a2] = at;
                            at += cnt[k2][a2];
                        }
            }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < OCC_K; k++)
    {
        if (it[k].x < 0) continue;   /* uniform */
        int p0 = 0, p1 = 0;
        for (int v = 0; v < w; v++)
        {
            p0 += wc[k][0][v];
            p1 += wc[k][1][v];
        }
        const int c = it[k].y * TPB + (int)threadIdx.x;
        if (n[k] > 0)
        {
            const int cat = lst[k][0] - 1;
            const int q = p0 + __popcll(occ[k] & below);
            P.occ_cell0[cat][run[k][0] + q] = make_int2(it[k].x, c);
            if (lst[k][1] >= 0) P.occ_cell1[cat][run[k][1] + q] = make_int2(it[k].x, c);
            P.occ_all[run[k][3] + q] = make_int2(it[k].x, c);
        }
        if ((gm[k] >> lane) & 1ull) P.occ_slot[run[k][2] + p1 + __popcll(gm[k] & below)] = make_int2(it[k].x, c / CPB);
    }
}

static void launch_occ(const BatchPlan &P, hipStream_t st)
{
    BatchMap mc = {P.descs, P.cellblk_items, nullptr, 0};
    hipLaunchKernelGGL(kb_occ, dim3((P.n_cellblk + OCC_K - 1) / OCC_K), dim3(TPB), 0, st, mc, P);
}

/* outside a step (a batch's set-up): every list length zeroed, then the lists */
hipError_t launchb_occ(const BatchPlan &P, hipStream_t st)
{
    if (P.nsys == 0 || !P.compact) return hipSuccess;
    hipLaunchKernelGGL(kb_occ_clear, dim3(1), dim3(64), 0, st, P.occ_n, 0xffffu);
    launch_occ(P, st);
    return hipGetLastError();
}

/* a batch launch over the slot groups: every group, or the occupied ones */
static inline void slot_map(const BatchPlan &P, BatchMap &m, dim3 &g)
{
    if (P.compact)
    {
        m = {P.descs, P.occ_slot, P.occ_n, 0};
        g = dim3((unsigned)P.g_slot);
    }
    else
    {
        m = {P.descs, P.slot_items, nullptr, 0};
        g = dim3((unsigned)P.n_slot);
    }
}

/* ------------------------------------------------------ batched launchers */

hipError_t launchb_vv_start(const BatchPlan &P, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    BatchMap ms, mc = {P.descs, P.cellblk_items, nullptr, 0};
    dim3 gs, gc(P.n_cellblk);
    slot_map(P, ms, gs);
    if (P.compact)
    {
        const BatchMap ma = {P.descs, P.occ_all, P.occ_n + 10, 0};
        hipLaunchKernelGGL(kb_vv_start_cells, dim3((unsigned)P.g_all), dim3(TPB), 0, st, ma, P.occ_n + 9);
    }
    else
        hipLaunchKernelGGL(kb_vv_start, gs, dim3(TPB), 0, st, ms);
    if (P.compact)
    {
        hipLaunchKernelGGL(kb_touch, dim3((P.n_cellblk + TPB - 1) / TPB), dim3(TPB), 0, st, P.descs,
                           P.cellblk_items, P.n_cellblk, P.occ_bin, P.occ_n);
        mc = {P.descs, P.occ_bin, P.occ_n + 9, 0};
        gc = dim3((unsigned)P.g_bin);
    }
    hipLaunchKernelGGL(kb_bin_rank, gc, dim3(TPB), 0, st, mc);
    hipLaunchKernelGGL(kb_bin_apply, gc, dim3(TPB), 0, st, mc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (P.compact) launch_occ(P, st);   /* the lists of the atoms' new cells (kb_touch zeroed their lengths) */
    return hipGetLastError();
}

hipError_t launchb_vv_finish(const BatchPlan &P, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    BatchMap ms;
    dim3 gs;
    slot_map(P, ms, gs);
    hipLaunchKernelGGL(kb_vv_finish, gs, dim3(TPB), 0, st, ms);
    hipLaunchKernelGGL(kb_reduce_thermo, dim3(P.nsys), dim3(1024), 0, st, P.descs);
    return hipGetLastError();
}

hipError_t launchb_thermo(const BatchPlan &P, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    BatchMap ms = {P.descs, P.slot_items, nullptr, 0};
    hipLaunchKernelGGL(kb_thermo, dim3(P.n_slot), dim3(TPB), 0, st, ms);
    hipLaunchKernelGGL(kb_reduce_thermo, dim3(P.nsys), dim3(1024), 0, st, P.descs);
    return hipGetLastError();
}

hipError_t launchb_reduce_pe(const BatchPlan &P, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    BatchMap mp = {P.descs, P.pe_items, nullptr, 0};
    hipLaunchKernelGGL(kb_pe_part, dim3(P.n_pe), dim3(TPB), 0, st, mp);
    hipLaunchKernelGGL(kb_pe_final, dim3(P.nsys), dim3(1024), 0, st, P.descs);
    return hipGetLastError();
}

hipError_t launchb_collect(const BatchPlan &P, hipStream_t st, unsigned *status, double *thermo)
{
    if (P.nsys == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_collect, dim3(P.nsys), dim3(64), 0, st, P.descs, status, thermo);
    return hipGetLastError();
}
