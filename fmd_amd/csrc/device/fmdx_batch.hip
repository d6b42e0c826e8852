/*
 * fmdx_batch.hip -- many independent systems advanced together.
 *
 * The reference runs one system per fmd_t.  A batch keeps the device state of
 * every member system where it already lives (each member is an uploaded
 * fmdx_system_t) and replaces the per-system launches by one launch per
 * kernel over a work list of (system, logical block) items:
 *   slot items     (system, group of 16 cells)   integrators, relative positions
 *   cell-block     (system, block of 256 cells)  exact-order re-binning
 *   cell items     (system, cell)                pair kernels, one list per
 *                                                pair category and pass
 *   energy items   (system, partial block)       fixed-order energy sums
 * Each item runs exactly the code and summation order of the single-system
 * path, so a member's trajectory does not depend on what else is in the
 * batch.  The lists are ordered so that the round-robin workgroup dispatch
 * gives each XCD a contiguous run of items (neighbour cells share its L2).
 * A whole velocity-Verlet step is captured once in a HIP graph and replayed,
 * which leaves one host launch per step: two graphs, one that also sums the
 * pair energies for the steps whose status is read, one without them.
 *
 * Between steps the host reads every member's status word with one copy
 * (kb_collect) to grow slot capacities before they can overflow and to
 * report members whose device error word is set.
 *
 * Members that are mostly vacuum (clusters in an open box) are stepped over
 * their occupied cells only: after each re-binning the device compacts the
 * cell, group and block lists (kb_occ, kb_touch in k_step.hip), and the
 * captured step walks them with grids the host sizes from the list lengths
 * it reads with the status words (fit_grids).  Empty cells and groups get the
 * words the skipped kernels would have written, so every result is the same.
 */
#include <algorithm>
#include <vector>
#include "fmdx_sys.h"
#include "fmdx_tuned.h"

struct fmdx_batch
{
    int dev;
    hipStream_t st;
    int precision;
    std::vector<fmdx_system_t *> sys;
    KDesc *descs;
    int descs_cap;
    int2 *items;
    size_t items_cap;
    char *occ_mem;             /* compacted lists: lengths, member categories, the lists */
    size_t occ_cap;            /* list entries */
    int occ_sys_cap;
    unsigned *dstatus;
    double *dthermo;
    int status_cap;
    std::vector<unsigned> hstatus;
    std::vector<double> hthermo;
    int hocc[16];              /* lengths of the compacted lists, as last read */
    int compact_mode;          /* FMDX_COMPACT_* */
    int ncells;                /* cells of all members */
    BatchPlan plan;
    int dirty;
    int use_graph;
    hipGraph_t graph[2];       /* [energy]: the step without / with the pair energies */
    hipGraphExec_t gexec[2];
    int graph_ok;
    int check_every;
};

/* item b of a list of n runs on XCD b % 8: give XCD x the contiguous run of
   list positions x*q + min(x, r) ... (q = n / 8, r = n % 8) */
static void xcd_order(std::vector<int2> &v)
{
    const size_t n = v.size();
    if (n < 16) return;
    std::vector<int2> o(n);
    const size_t q = n / 8, r = n % 8;
    for (size_t b = 0; b < n; b++)
    {
        const size_t x = b % 8, i = b / 8;
        o[b] = v[x * q + std::min(x, r) + i];
    }
    v.swap(o);
}

static void graph_drop(fmdx_batch_t *b)
{
    if (b->graph_ok)
        for (int k = 0; k < 2; k++)
        {
            hipGraphExecDestroy(b->gexec[k]);
            hipGraphDestroy(b->graph[k]);
        }
    b->graph_ok = 0;
}

static int upload_descs(fmdx_batch_t *b)
{
    const int n = (int)b->sys.size();
    std::vector<KDesc> h(n > 0 ? n : 1);
    for (int i = 0; i < n; i++)
    {
        h[i] = b->sys[i]->D;
        h[i].xq_keep = b->precision == FMDX_PREC_MIXED;
    }
    if (n) CK(hipMemcpyAsync(b->descs, h.data(), sizeof(KDesc) * n, hipMemcpyHostToDevice, b->st));
    /* the copy reads h, which ends with this call */
    CK(hipStreamSynchronize(b->st));
    return FMDX_OK;
}

/* Whether the members' occupied cells, as last counted, are few enough for the compacted
   lists to pay: below 60 % of their cells.  Fuller batches take the full lists, whose
   launches skip the list lengths and the compaction itself. */
static bool sparse(const fmdx_batch_t *b)
{
    const BatchPlan &P = b->plan;
    long occ = 0, all = 0;
    for (int c = 0; c < 4; c++)
    {
        occ += b->hocc[1 + c];
        all += P.n_cell0[c];
    }
    return occ * 10 < all * 6;
}

/* Grids of the compacted lists from their lengths: room for a quarter more and 64 items
   (none when tight), at most one workgroup per item of the full list.  A grid is resized
   when its list outgrows it or needs less than half of it; returns whether any grid
   changed (the captured step then no longer matches). */
static int fit_grids(fmdx_batch_t *b, int force)
{
    BatchPlan &P = b->plan;
    if (!P.compact) return 0;
    const int tight = b->compact_mode == FMDX_COMPACT_TIGHT;
    auto fit = [&](int &g, int n, int full) {
        const int cap = std::max(8, (full + 7) / 8 * 8);
        const int room = tight ? 0 : n / 4 + 64;
        const int want = std::min(cap, std::max(8, (n + room + 7) / 8 * 8));
        if (want != g && (force || n > g || want < g / 2))
        {
            g = want;
            return 1;
        }
        return 0;
    };
    int changed = fit(P.g_slot, b->hocc[0], P.n_slot);
    changed |= fit(P.g_bin, force ? P.n_cellblk : b->hocc[9], P.n_cellblk);   /* no count before a step */
    changed |= fit(P.g_all, (b->hocc[10] + CPB - 1) / CPB, (b->ncells + CPB - 1) / CPB);   /* CPB cells a pass */
    for (int c = 0; c < 4; c++)
    {
        changed |= fit(P.g_cell0[c], b->hocc[1 + c], P.n_cell0[c]);
        changed |= fit(P.g_cell1[c], b->hocc[5 + c], P.n_cell1[c]);
    }
    return changed;
}

static int rebuild(fmdx_batch_t *b)
{
    const int n = (int)b->sys.size();
    graph_drop(b);
    if (n > b->descs_cap)
    {
        dfree(b->descs);
        int cap = std::max(n, 2 * b->descs_cap);
        if (dalloc(&b->descs, cap)) return FMDX_ENOMEM;
        b->descs_cap = cap;
    }
    if (n > b->status_cap)
    {
        dfree(b->dstatus);
        dfree(b->dthermo);
        int cap = std::max(n, 2 * b->status_cap);
        if (dalloc(&b->dstatus, (size_t)cap * 8) || dalloc(&b->dthermo, (size_t)cap * 16)) return FMDX_ENOMEM;
        b->status_cap = cap;
    }
    std::vector<int2> slot, cblk, pe, c0[4], c1[4];
    int maxk = 1;
    for (int i = 0; i < n; i++)
    {
        const KDesc &D = b->sys[i]->D;
        const int cat = force_category(D);
        maxk = std::max(maxk, D.nkinds);
        for (int g = 0; g < (D.ncell + CPB - 1) / CPB; g++) slot.push_back(make_int2(i, g));
        for (int g = 0; g < (D.ncell + TPB - 1) / TPB; g++) cblk.push_back(make_int2(i, g));
        for (int g = 0; g < pe_blocks(D.ncell); g++) pe.push_back(make_int2(i, g));
        for (int c = 0; c < D.ncell; c++)
        {
            c0[cat].push_back(make_int2(i, c));
            if (D.has_eam) c1[cat].push_back(make_int2(i, c));
        }
    }
    std::vector<int2> *lists[11] = {&slot, &cblk, &pe, &c0[0], &c0[1], &c0[2], &c0[3], &c1[0], &c1[1], &c1[2], &c1[3]};
    size_t total = 0;
    for (auto *l : lists)
    {
        xcd_order(*l);
        total += l->size();
    }
    if (total > b->items_cap)
    {
        dfree(b->items);
        if (dalloc(&b->items, total)) return FMDX_ENOMEM;
        b->items_cap = total;
    }
    std::vector<int2> all;
    all.reserve(total);
    const int2 *ptr[11];
    for (int k = 0; k < 11; k++)
    {
        ptr[k] = b->items + all.size();
        all.insert(all.end(), lists[k]->begin(), lists[k]->end());
    }
    if (total) CK(hipMemcpyAsync(b->items, all.data(), sizeof(int2) * total, hipMemcpyHostToDevice, b->st));
    /* the compacted lists hold at most every slot group and cell of the full ones */
    BatchPlan &P = b->plan;
    P.compact = FMDX_COMPACT && n > 0 && b->compact_mode != FMDX_COMPACT_OFF;
    size_t nocc = slot.size() + cblk.size(), ncells = 0;
    for (int c = 0; c < 4; c++)
    {
        nocc += 2 * c0[c].size() + c1[c].size();   /* the per-category lists and the list of all cells */
        ncells += c0[c].size();
    }
    auto occ_layout = [&](Carve &cv, size_t cap, size_t nsys) {
        cv(&P.occ_n, 16);
        cv((int **)&P.sys_cat, nsys);
        cv(&P.occ_slot, cap);   /* the whole list area */
    };
    if (P.compact && (nocc > b->occ_cap || n > b->occ_sys_cap))
    {
        const size_t cap = std::max(nocc, 2 * b->occ_cap);
        const int scap = std::max(n, 2 * b->occ_sys_cap);
        dfree(b->occ_mem);
        b->occ_cap = 0;
        b->occ_sys_cap = 0;
        if (carve_alloc(&b->occ_mem, [&](Carve &cv) { occ_layout(cv, cap, scap); })) return FMDX_ENOMEM;
        b->occ_cap = cap;
        b->occ_sys_cap = scap;
    }
    /* source of an asynchronous copy below: it lives until the stream synchronisation
       in upload_descs (or before an early return), like `all` */
    std::vector<int> cats;
    if (P.compact)
    {
        /* one carve of the whole list area: slot groups first, then each category's lists */
        int2 *at = P.occ_slot + slot.size();
        for (int c = 0; c < 4; c++)
        {
            P.occ_cell0[c] = at;
            at += c0[c].size();
            P.occ_cell1[c] = at;
            at += c1[c].size();
        }
        P.occ_bin = at;
        P.occ_all = at + cblk.size();
        b->ncells = (int)ncells;
        cats.resize(n);
        for (int i = 0; i < n; i++) cats[i] = force_category(b->sys[i]->D);
        CK(hipMemcpyAsync((void *)P.sys_cat, cats.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->st));
    }
    /* members whose force passes evaluate each pair once need their partner sums */
    for (int i = 0; i < n; i++)
    {
        const int e = ensure_n3(b->sys[i], b->precision);
        if (e)
        {
            hipStreamSynchronize(b->st);   /* the copies from `all` and `cats` end first */
            return e;
        }
    }
    int r = upload_descs(b);
    if (r) return r;
    P.descs = b->descs;
    P.nsys = n;
    P.slot_items = ptr[0]; P.n_slot = (int)slot.size();
    P.cellblk_items = ptr[1]; P.n_cellblk = (int)cblk.size();
    P.pe_items = ptr[2]; P.n_pe = (int)pe.size();
    for (int c = 0; c < 4; c++)
    {
        P.cell0[c] = ptr[3 + c]; P.n_cell0[c] = (int)c0[c].size();
        P.cell1[c] = ptr[7 + c]; P.n_cell1[c] = (int)c1[c].size();
    }
    P.max_nkinds = maxk;
    P.n3 = b->precision == FMDX_PREC_MIXED;   /* every LJ and Morse member holds partner sums */
    b->hstatus.assign((size_t)std::max(n, 1) * 8, 0);
    b->hthermo.assign((size_t)std::max(n, 1) * 16, 0.0);
    /* the first step's start and kick walk the lists of the members' current cells */
    CK(launchb_occ(P, b->st));
    if (P.compact) CK(hipMemcpyAsync(b->hocc, P.occ_n, sizeof(int) * 11, hipMemcpyDeviceToHost, b->st));
    CK(hipStreamSynchronize(b->st));
    if (b->compact_mode == FMDX_COMPACT_AUTO) P.compact = P.compact && sparse(b);
    fit_grids(b, 1);
    b->dirty = 0;
    return FMDX_OK;
}

/* one step; `energy` also sums the pair energies, which only a status read needs */
static hipError_t enqueue_step(fmdx_batch_t *b, int energy)
{
    hipError_t e = launchb_vv_start(b->plan, b->st);
    /* mixed: the start step and the re-binning left every member's FP32 cell-relative
       positions in place (the descriptors carry xq_keep), so the force passes need no
       separate pass forming them */
    if (e == hipSuccess) e = launchb_forces(b->plan, b->st, b->precision, energy, b->precision != FMDX_PREC_MIXED);
    if (e == hipSuccess) e = launchb_vv_finish(b->plan, b->st);
    return e;
}

static int capture(fmdx_batch_t *b)
{
    graph_drop(b);
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t x[2] = {nullptr, nullptr};
    int ok = 1;
    for (int k = 0; k < 2 && ok; k++)
    {
        if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        {
            ok = 0;
            break;
        }
        const hipError_t e = enqueue_step(b, k);
        const hipError_t e2 = hipStreamEndCapture(b->st, &g[k]);
        ok = e == hipSuccess && e2 == hipSuccess && hipGraphInstantiate(&x[k], g[k], nullptr, nullptr, 0) == hipSuccess;
    }
    if (!ok)
    {
        for (int k = 0; k < 2; k++)
        {
            if (x[k]) hipGraphExecDestroy(x[k]);
            if (g[k]) hipGraphDestroy(g[k]);
        }
        return FMDX_EHIP;
    }
    for (int k = 0; k < 2; k++)
    {
        b->graph[k] = g[k];
        b->gexec[k] = x[k];
    }
    b->graph_ok = 1;
    return FMDX_OK;
}

/* one copy of every member's status words and thermo record */
static int collect(fmdx_batch_t *b)
{
    const int n = (int)b->sys.size();
    if (!n) return FMDX_OK;
    CK(launchb_collect(b->plan, b->st, b->dstatus, b->dthermo));
    CK(hipMemcpyAsync(b->hstatus.data(), b->dstatus, sizeof(unsigned) * 8 * n, hipMemcpyDeviceToHost, b->st));
    CK(hipMemcpyAsync(b->hthermo.data(), b->dthermo, sizeof(double) * 16 * n, hipMemcpyDeviceToHost, b->st));
    if (b->plan.compact)
        CK(hipMemcpyAsync(b->hocc, b->plan.occ_n, sizeof(int) * 11, hipMemcpyDeviceToHost, b->st));
    CK(hipStreamSynchronize(b->st));
    if (b->plan.compact && b->compact_mode == FMDX_COMPACT_AUTO && !sparse(b))
    {
        b->plan.compact = 0;   /* the members filled their cells: the full lists from here on */
        graph_drop(b);
    }
    else if (fit_grids(b, 0))
        graph_drop(b);   /* captured again before the next step */
    return FMDX_OK;
}

/* grow the slot capacity of members whose fullest cell nears it */
static int grow(fmdx_batch_t *b)
{
    int changed = 0;
    for (size_t i = 0; i < b->sys.size(); i++)
    {
        fmdx_system_t *s = b->sys[i];
        const unsigned maxc = b->hstatus[i * 8 + 2];
        const unsigned cap = (unsigned)s->D.cap;
        if (cap >= 256 || maxc + 8 + maxc / 4 <= cap) continue;
        unsigned nc = maxc + maxc / 2 + 8;
        nc = (nc + 31) & ~31u;   /* multiples of 32 slots: each cell starts on a cache line */
        if (nc > 256) nc = 256;
        /* the batch itself swaps the member's slot arrays: its descriptors are refreshed below */
        s->member = 0;
        int r = fmdx_set_capacity(s, nc);
        s->member = 1;
        if (r) return r;
        if ((r = ensure_n3(s, b->precision))) return r;   /* the new layout's partner sums */
        changed = 1;
    }
    if (changed)
    {
        /* new slot arrays: the captured graph reads descriptors from device
           memory, so refreshing them is enough */
        int r = upload_descs(b);
        if (r) return r;
        CK(hipStreamSynchronize(b->st));
    }
    return FMDX_OK;
}

extern "C" int fmdx_batch_create(fmdx_batch_t **out, int precision)
{
    if (!out || (precision != FMDX_PREC_FP64 && precision != FMDX_PREC_MIXED)) return FMDX_EARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return FMDX_ENODEV;
    fmdx_batch_t *b = new fmdx_batch_t();
    CK(hipGetDevice(&b->dev));
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
    {
        delete b;
        return FMDX_EHIP;
    }
    b->precision = precision;
    b->use_graph = 1;
    b->check_every = 8;
    b->compact_mode = FMDX_COMPACT_AUTO;
    b->dirty = 1;
    *out = b;
    return FMDX_OK;
}

extern "C" void fmdx_batch_destroy(fmdx_batch_t *b)
{
    if (!b) return;
    hipStreamSynchronize(b->st);
    for (auto *s : b->sys) s->member = 0;
    graph_drop(b);
    dfree(b->descs);
    dfree(b->items);
    dfree(b->occ_mem);
    dfree(b->dstatus);
    dfree(b->dthermo);
    hipStreamDestroy(b->st);
    delete b;
}

/* Join a batch.  The member's forces and thermo record are brought up to date
   on its own stream first, as the reference's integrate loop does before its
   first step (integrators.c:232). */
extern "C" int fmdx_batch_add(fmdx_batch_t *b, fmdx_system_t *s, double dt, int thermostat, double tau, double T0)
{
    if (!b || !s) return FMDX_EARG;
    if (!s->uploaded) return FMDX_ESTATE;
    if (s->ttm_on) return FMDX_EARG;   /* TTM members are not batched */
    for (auto *t : b->sys)
        if (t == s) return FMDX_EARG;
    if (s->member) return FMDX_ESTATE;   /* already in another batch */
    int r = fmdx_set_precision(s, b->precision);
    if (r) return r;
    r = fmdx_update_forces(s, -1, 0);
    if (r) return r;
    r = fmdx_thermo(s, -1);
    if (r) return r;
    CK(hipStreamSynchronize(s->st));
    set_step(s->D, s->mass, -1, dt, thermostat, tau, T0);
    s->D.energy = 1;   /* its own force calls sum the energies; batched steps do on status steps */
    s->member = 1;
    b->sys.push_back(s);
    b->dirty = 1;
    return FMDX_OK;
}

extern "C" int fmdx_batch_remove(fmdx_batch_t *b, fmdx_system_t *s)
{
    if (!b || !s) return FMDX_EARG;
    auto it = std::find(b->sys.begin(), b->sys.end(), s);
    if (it == b->sys.end()) return FMDX_EARG;
    CK(hipStreamSynchronize(b->st));
    s->member = 0;
    b->sys.erase(it);
    b->dirty = 1;
    return FMDX_OK;
}

extern "C" int fmdx_batch_size(fmdx_batch_t *b, int *n)
{
    if (!b || !n) return FMDX_EARG;
    *n = (int)b->sys.size();
    return FMDX_OK;
}

extern "C" int fmdx_batch_use_graph(fmdx_batch_t *b, int on)
{
    if (!b) return FMDX_EARG;
    b->use_graph = on ? 1 : 0;
    if (!on) graph_drop(b);
    return FMDX_OK;
}

extern "C" int fmdx_batch_set_check_interval(fmdx_batch_t *b, int steps)
{
    if (!b || steps < 1) return FMDX_EARG;
    b->check_every = steps;
    return FMDX_OK;
}

extern "C" int fmdx_batch_set_compaction(fmdx_batch_t *b, int mode)
{
    if (!b || mode < FMDX_COMPACT_OFF || mode > FMDX_COMPACT_TIGHT) return FMDX_EARG;
    if (mode != b->compact_mode)
    {
        b->compact_mode = mode;
        b->dirty = 1;   /* lists, grids and the captured step are laid out again */
    }
    return FMDX_OK;
}

/* Velocity-Verlet steps of every member (VelocityVerlet_takeOneStep,
   integrators.c:181-211, with each member's own dt and thermostat). */
extern "C" int fmdx_batch_step(fmdx_batch_t *b, int nsteps)
{
    if (!b || nsteps < 0) return FMDX_EARG;
    int r;
    if (b->dirty && (r = rebuild(b))) return r;
    if (b->sys.empty()) return FMDX_OK;
    for (int k = 0; k < nsteps; k++)
    {
        /* the steps after which the status is read sum the pair energies */
        const int energy = (k + 1) % b->check_every == 0 || k + 1 == nsteps;
        if (b->use_graph && !b->graph_ok && (r = capture(b))) return r;
        if (b->use_graph) CK(hipGraphLaunch(b->gexec[energy], b->st));
        else CK(enqueue_step(b, energy));
        if (energy)
        {
            if ((r = collect(b))) return r;
            if ((r = grow(b))) return r;
        }
    }
    return FMDX_OK;
}

/* thermo record and status of every member, in join order */
extern "C" int fmdx_batch_status(fmdx_batch_t *b, fmdx_thermo_t *out)
{
    if (!b || !out) return FMDX_EARG;
    int r;
    if (b->dirty && (r = rebuild(b))) return r;
    if ((r = collect(b))) return r;
    for (size_t i = 0; i < b->sys.size(); i++)
    {
        const double *th = &b->hthermo[i * 16];
        const unsigned *ct = &b->hstatus[i * 8];
        fmdx_thermo_t &o = out[i];
        o.pe = th[TH_PE];
        o.embed = th[TH_EMB];
        o.mv2 = th[TH_MV2];
        o.ke = th[TH_KE];
        o.temperature = th[TH_T];
        o.momentum[0] = th[TH_PX];
        o.momentum[1] = th[TH_PY];
        o.momentum[2] = th[TH_PZ];
        o.n_active = th[TH_N];
        o.natoms = b->sys[i]->natoms_uploaded - ct[1];
        o.max_cell_count = ct[2];
        o.err = ct[3];
        o.cap = (unsigned)b->sys[i]->D.cap;
        o.bad_force_cell = ct[4];
        o.bad_position_cell = ct[5];
        o.replay_mismatch = ct[6];
        o.exact_unmatched = ct[7];
        o.audit_atoms = o.audit_max_cell = o.audit_nonfinite = 0;
    }
    return FMDX_OK;
}

extern "C" int fmdx_batch_sync(fmdx_batch_t *b)
{
    if (!b) return FMDX_EARG;
    CK(hipStreamSynchronize(b->st));
    return FMDX_OK;
}
