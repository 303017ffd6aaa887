// kmcg_otf.hip -- the otf backend: per-site rates from the local environment.
//
// Included by kmcg_engine.hip (one translation unit: it shares the error
// helpers and the uniform stream).  The reference's otf step loop
// (kmos/fortran_src/base_otf.f90, the generated proclist of
// kmos/io/__init__.py:2657-3723) keeps, per process, a list of the cells
// where it can run (avail_sites) and each entry's own rate
// (rates_matrix); every step re-sums every list in list order
// (update_accum_rate, base_otf.f90:687-717), picks a process by interval
// search over the process sums and a cell by interval search over the
// running sum of the chosen list (determine_procsite, 1213-1278), then
// runs the process: deletes, lattice writes, rate updates of processes
// with a bystander on a changed site, and adds with the new rates.
//
// Layout.  One wavefront per replica, lane q owns process q (at most 64
// processes).  The lattice sits in LDS as one byte per site.  In HBM, per
// replica and process, four rows of C entries (C = cells): the list
// (cell + 1, 0 = empty), the inverse index (list position + 1 by cell),
// the entry rates and their running sum in list order.  The running sum is
// what makes the step exact and cheap: the reference's full re-summation
// in list order equals the stored running sum as long as every entry from
// the first changed position on is re-added, so a step re-adds only the
// tail behind the smallest list position it changed (per owner lane,
// sequential, in the same order as the reference's loop).  The site
// search reads that running sum: 64-ary, two or three dependent rounds.
//
// Rates.  gr_<proc>(cell) depends only on the bystander counts; the host
// evaluates each process's expression for every count combination once per
// parameter set (Fortran kinds: see kmos_amd/otf.py) and the kernel reads
// table[base + sum of per-bystander strides].
namespace {

constexpr int OTF_MAX_P = 64;
constexpr int OTF_MAX_V = 65536;     // sites per replica (one byte each in LDS)
constexpr int OTF_LDS = 65536;       // LDS bytes a replica's wavefront may use

// The device program, compact int32 words (built by otf_build):
//   [0, spuck)        site defaults
//   owner [P][P]      run_proc_<p>'s ops for process q: off | n_del << 16 | n_rest << 24
//   writes [P]        off | count << 16 -> 2 words per write: packed site, old | new << 16
//   byst [P]          off | count << 16 -> per bystander: packed site, then S strides
//   tbase [P]         first table entry of process q
//   touch [P]         off | count << 16 -> touchup_cell's add ops of process q
// op word:   kind (2 bits: 0 del, 2 update, 3 add) | dx, dy, dz (8 bits each) << 2, 10, 18
//            | ncond << 26 (adds)
// cond word: dx, dy, dz (7 bits each) | site - 1 (5 bits) << 21 | species << 26
// site word: dx, dy, dz (8 bits each) | site - 1 << 24
struct OtfParams {
    int P, S, spuck, T, W;
    const int *prog;                 // global copy of the program (W words)
    int off_owner, off_writes, off_byst, off_tbase, off_touch;
    int tab_lds;                     // 1: the replica's rate table is staged in LDS too
    const int *dims;                 // [n][4]: X, Y, Z, C
    const long long *lat_off;        // [n]: first site of the replica in `lat`
    const long long *list_off;       // [n]: first entry of the replica's rows
    signed char *lat;
    int *sites, *back;
    double *rm, *pre;
    int *nsites;                     // [n][P]
    long long *procstat;             // [n][P]
    double *integ, *accum, *runsum;  // [n][P]
    double *table;                   // [n][T]
    double *kmc_time, *kmc_time_step;
    long long *kmc_step;
    unsigned long long *ctr;
    const unsigned *seed, *stream;
    int *status;
    const int *rep_ids;
    long long steps;
    int mode;
};

__device__ __forceinline__ int otf_wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
__device__ __forceinline__ int sx8(int w, int sh) { return (int)(signed char)((w >> sh) & 0xFF); }
__device__ __forceinline__ int sx7(int w, int sh) { return (((w >> sh) & 0x7F) ^ 0x40) - 0x40; }

struct OtfCell {
    int X, Y, Z, C;
    __device__ __forceinline__ int idx(int x, int y, int z) const { return x + X * (y + Y * z); }
    __device__ __forceinline__ void xyz(int c, int &x, int &y, int &z) const
    {
        x = c % X;
        const int yz = c / X;
        y = yz % Y;
        z = yz / Y;
    }
    // the cell at (x, y, z) + (dx, dy, dz), each offset at most one period
    __device__ __forceinline__ int at(int x, int y, int z, int dx, int dy, int dz) const
    {
        return idx(otf_wrap(x + dx, X), otf_wrap(y + dy, Y), otf_wrap(z + dz, Z));
    }
};

// gr_<q>(cell): the table entry of the bystander counts around (x, y, z)
__device__ __forceinline__ double otf_gr(const int *prog, const OtfParams &p, const OtfCell &g,
                                         const signed char *slat, const double *tab, int q, int x, int y, int z)
{
    const int bt = prog[p.off_byst + q];
    const int *b = prog + (bt & 0xFFFF);
    const int nb = bt >> 16;
    int idx = 0;
    for (int k = 0; k < nb; ++k, b += 1 + p.S) {
        const int w = b[0];
        const int c = g.at(x, y, z, sx8(w, 0), sx8(w, 8), sx8(w, 16));
        const int sp = slat[p.spuck * c + ((w >> 24) & 0xFF)];
        if (sp >= 0 && sp < p.S)
            idx += b[1 + sp];
    }
    return tab[prog[p.off_tbase + q] + idx];
}

// del_proc (base_otf.f90:221-282) on the owner lane's rows: the last entry
// moves into the freed position.  Returns the first list position whose
// rate changed (the freed one, or the new end when the moved entry carries
// the same rate, so that the running sums below it stand), -1 when the
// inverse index points outside the list.
__device__ __forceinline__ int otf_del(int *srow, int *brow, double *rrow, int &n, double &run, int c)
{
    const int pos = brow[c] - 1;
    const int last = n - 1;
    if (pos < 0 || pos > last)
        return -1;
    const double gone = rrow[pos];
    run = run - gone;
    int changed = last;
    if (pos < last) {
        const int mv = srow[last];
        const double mr = rrow[last];
        srow[pos] = mv;
        rrow[pos] = mr;
        brow[mv - 1] = pos + 1;
        if (mr != gone)
            changed = pos;
    }
    srow[last] = 0;
    rrow[last] = 0.0;
    brow[c] = 0;
    n = last;
    return changed;
}

// the running sum of a row from position `from` on (update_accum_rate's
// order: left to right), one lane alone; returns the row's total
__device__ __forceinline__ double otf_resum(const double *rrow, double *prow, int from, int n)
{
    double a = from > 0 ? prow[from - 1] : 0.0;
    int j = from;
    for (; j + 8 <= n; j += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            v[k] = rrow[j + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a = a + v[k];
            prow[j + k] = a;
        }
    }
    for (; j < n; ++j) {
        a = a + rrow[j];
        prow[j] = a;
    }
    return a;
}

// lane l receives v from lane l-1 of its row of 16 (DPP row_shr:1); the
// first lane of each row receives 0
__device__ __forceinline__ double row_shr1z(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x111, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x111, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The pending list with the most entries to re-add (uniform; ties to the
// lowest lane), taken off `pend`; -1 when none is left.
__device__ __forceinline__ int otf_take_longest(unsigned long long &pend, int len)
{
    int best = -1, most = -1;
    for (unsigned long long m = pend; m; m &= m - 1) {
        const int l = __ffsll((long long)m) - 1;
        const int v = readlane_i(len, l);
        if (v > most) {
            most = v;
            best = l;
        }
    }
    if (best >= 0)
        pend &= ~(1ull << best);
    return best;
}

// The running sums of every list a step changed, re-added from its first
// changed position to its end, four lists at a time (the longest four
// first, so that lists of like length share a pass): each row of 16 lanes
// walks one list in blocks of 16 entries (one coalesced read), and within
// a block the sum is carried left to right through the row by 15 shifted
// adds, so lane i ends with ((carry + r0) + r1) + ... + ri, the rounding of
// update_accum_rate's loop.  `dirty` and `n` are per owner lane.
__device__ __forceinline__ void otf_resum_rows(const double *rm, double *pre, long long base, int C, int lane,
                                               bool want, int dirty, int n)
{
    unsigned long long dm = __ballot(want && dirty < n);
    const int row = lane >> 4, i = lane & 15;
    const int mylen = (want && dirty < n) ? n - dirty : 0;
    while (dm) {
        int qs[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            qs[k] = otf_take_longest(dm, mylen);
        int d4[4], n4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            d4[k] = qs[k] >= 0 ? readlane_i(dirty, qs[k]) : 0;
            n4[k] = qs[k] >= 0 ? readlane_i(n, qs[k]) : 0;
        }
        const int q = row == 0 ? qs[0] : (row == 1 ? qs[1] : (row == 2 ? qs[2] : qs[3]));
        const int d = row == 0 ? d4[0] : (row == 1 ? d4[1] : (row == 2 ? d4[2] : d4[3]));
        const int len = (row == 0 ? n4[0] : (row == 1 ? n4[1] : (row == 2 ? n4[2] : n4[3]))) - d;
        const int maxlen = max(max(n4[0] - d4[0], n4[1] - d4[1]), max(n4[2] - d4[2], n4[3] - d4[3]));
        const double *rr = rm + base + (long long)(q < 0 ? 0 : q) * C;
        double *pr = pre + base + (long long)(q < 0 ? 0 : q) * C;
        double carry = (q >= 0 && d > 0) ? pr[d - 1] : 0.0;
        // eight blocks per trip, the next eight read while these are summed
        constexpr int NB = 8;
        double v[NB], vn[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
            v[b] = (16 * b + i < len) ? rr[d + 16 * b + i] : 0.0;
        for (int off = 0; off < maxlen; off += 16 * NB) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int jn = off + 16 * (NB + b) + i;
                vn[b] = (jn < len) ? rr[d + jn] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = off + 16 * b;
                if (j >= maxlen)
                    break;
                const double ce = (i == 0) ? carry + v[b] : v[b];
                double a = ce;
#pragma unroll
                for (int k = 1; k < 16; ++k)
                    a = row_shr1z(a) + ce;
                if (j + i < len)
                    pr[d + j + i] = a;
                const double c0 = readlane_d(a, 15), c1 = readlane_d(a, 31), c2 = readlane_d(a, 47),
                             c3 = readlane_d(a, 63);
                carry = row == 0 ? c0 : (row == 1 ? c1 : (row == 2 ? c2 : c3));
            }
#pragma unroll
            for (int b = 0; b < NB; ++b)
                v[b] = vn[b];
        }
    }
}

__device__ __forceinline__ bool otf_conds(const OtfParams &p, const OtfCell &g, const signed char *slat,
                                          const int *cd, int ncond, int x, int y, int z)
{
    for (int j = 0; j < ncond; ++j) {
        const int w = cd[j];
        const int c = g.at(x, y, z, sx7(w, 0), sx7(w, 7), sx7(w, 14));
        if (slat[p.spuck * c + ((w >> 21) & 0x1F)] != ((w >> 26) & 0x3F))
            return false;
    }
    return true;
}

// MODE 0: initialize_state (default species everywhere, touchup_cell over
// the cells with x fastest); MODE 1: _adjust_database after a new lattice
// (touchup_cell with x outermost: delete the cell's entry, then add it if
// the conditions hold); MODE 2: recalculate_rates_matrix (every listed
// entry takes its cell's rate under new parameters).  Each ends with the
// rows' running sums.  The program and tables are read from global memory
// (a one-off pass).
__global__ __launch_bounds__(64) void otf_setup_kernel(OtfParams p)
{
    extern __shared__ signed char slat[];
    const int r = p.rep_ids[blockIdx.x];
    const int lane = threadIdx.x;
    const OtfCell g{p.dims[4 * r], p.dims[4 * r + 1], p.dims[4 * r + 2], p.dims[4 * r + 3]};
    const int V = g.C * p.spuck;
    const int *prog = p.prog;
    signed char *glat = p.lat + p.lat_off[r];
    if (p.mode == 0) {
        for (int i = lane; i < V; i += 64)
            glat[i] = (signed char)prog[i % p.spuck];     // site defaults head the program
        __syncthreads();
    }
    for (int i = lane; i < V; i += 64)
        slat[i] = glat[i];
    __syncthreads();
    if (lane >= p.P)
        return;
    const int q = lane;
    const long long lo = p.list_off[r] + (long long)q * g.C;
    int *srow = p.sites + lo;
    int *brow = p.back + lo;
    double *rrow = p.rm + lo;
    double *prow = p.pre + lo;
    const double *tab = p.table + (long long)r * p.T;
    const int slot = r * p.P + q;
    int n = p.nsites[slot];
    double run = p.runsum[slot];
    if (p.mode == 0) {
        for (int c = 0; c < g.C; ++c) {
            srow[c] = 0;
            brow[c] = 0;
            rrow[c] = 0.0;
        }
        n = 0;
        run = 0.0;
    }
    if (p.mode == 2) {
        for (int j = 0; j < n; ++j) {
            if (srow[j] < 1 || srow[j] > g.C)
                continue;
            int x, y, z;
            g.xyz(srow[j] - 1, x, y, z);
            rrow[j] = otf_gr(prog, p, g, slat, tab, q, x, y, z);
        }
        // reaccumulate_rates_matrix (base_otf.f90:366-387): memory order
        run = 0.0;
        for (int j = 0; j < n; ++j)
            run = run + rrow[j];
    } else {
        const int tt = prog[p.off_touch + q];
        for (int k = 0; k < g.C; ++k) {
            int x, y, z;
            if (p.mode == 0) {
                g.xyz(k, x, y, z);
            } else {
                // x outermost, then y, then z
                z = k % g.Z;
                const int xy = k / g.Z;
                y = xy % g.Y;
                x = xy / g.Y;
            }
            const int c = g.idx(x, y, z);
            if (p.mode == 1 && brow[c] && otf_del(srow, brow, rrow, n, run, c) < 0) {
                p.status[r] = KMCG_REPLICA_CORRUPT;
                break;
            }
            const int *op = prog + (tt & 0xFFFF);
            for (int m = 0; m < (tt >> 16); ++m) {
                const int ncond = (op[0] >> 26) & 0x3F;
                if (otf_conds(p, g, slat, op + 1, ncond, x, y, z) && n < g.C) {
                    const double rate = otf_gr(prog, p, g, slat, tab, q, x, y, z);
                    srow[n] = c + 1;
                    brow[c] = n + 1;
                    rrow[n] = rate;
                    run = run + rate;
                    n += 1;
                }
                op += 1 + ncond;
            }
        }
    }
    otf_resum(rrow, prow, 0, n);
    p.nsites[slot] = n;
    p.runsum[slot] = run;
    if (p.mode == 0) {
        p.procstat[slot] = 0;
        p.integ[slot] = 0.0;
        p.accum[slot] = 0.0;
        if (q == 0) {
            p.kmc_time[r] = 0.0;
            p.kmc_time_step[r] = 0.0;
            p.kmc_step[r] = 0;
            p.ctr[r] = 0;
            p.status[r] = KMCG_REPLICA_OK;
        }
    }
}

// do_kmc_steps for `steps` steps (fewer when the replica stops).  LDS holds
// the program, the replica's rate table (TLDS) and its lattice.
template <bool TLDS>
__global__ __launch_bounds__(64) void otf_steps_kernel(OtfParams p)
{
    extern __shared__ __align__(16) int otf_smem[];
    const int r = p.rep_ids[blockIdx.x];
    const int lane = threadIdx.x;
    const int P = p.P;
    const OtfCell g{p.dims[4 * r], p.dims[4 * r + 1], p.dims[4 * r + 2], p.dims[4 * r + 3]};
    const int V = g.C * p.spuck;
    int *prog = otf_smem;
    const int W2 = (p.W + 1) & ~1;
    double *stab = reinterpret_cast<double *>(otf_smem + W2);
    signed char *slat = reinterpret_cast<signed char *>(otf_smem + W2 + (TLDS ? 2 * p.T : 0));
    signed char *glat = p.lat + p.lat_off[r];
    for (int i = lane; i < p.W; i += 64)
        prog[i] = p.prog[i];
    if (TLDS)
        for (int i = lane; i < p.T; i += 64)
            stab[i] = p.table[(long long)r * p.T + i];
    for (int i = lane; i < V; i += 64)
        slat[i] = glat[i];
    __syncthreads();
    const double *tab = TLDS ? stab : p.table + (long long)r * p.T;
    int st = p.status[r];
    if (st != KMCG_REPLICA_OK)
        return;
    const bool act = lane < P;
    const int q = act ? lane : 0;
    const long long base = p.list_off[r];
    const long long lo = base + (long long)q * g.C;
    int *srow = p.sites + lo;
    int *brow = p.back + lo;
    double *rrow = p.rm + lo;
    double *prow = p.pre + lo;
    const int slot = r * P + q;
    int n = act ? p.nsites[slot] : 0;
    double tot = (act && n > 0) ? prow[n - 1] : 0.0;
    double integ = act ? p.integ[slot] : 0.0;
    double run = act ? p.runsum[slot] : 0.0;
    double my_acc = act ? p.accum[slot] : 0.0;
    long long pstat = act ? p.procstat[slot] : 0;
    double t = p.kmc_time[r];
    double t_step = p.kmc_time_step[r];
    long long kstep = p.kmc_step[r];
    unsigned long long ctr = p.ctr[r];
    const unsigned seed = p.seed[r], stream = p.stream[r];
    for (long long s = 0; s < p.steps; ++s) {
        double u[3];
        step_uniforms(seed, stream, ctr, u);
        // update_accum_rate: accum(i) = accum(i-1) + sum of list i
        double acc = 0.0;
        for (int i = 0; i < P; ++i) {
            const double ti = readlane_d(tot, i);
            acc = i == 0 ? ti : acc + ti;
            if (lane == i)
                my_acc = acc;
        }
        run = tot;
        const double total = acc;
        if (!(total > 0.0)) {
            st = KMCG_REPLICA_ABSORBING;
            break;
        }
        ctr += 1;
        // update_clocks, update_integ_rate
        const double dt = -log(u[0]) / total;
        t_step = dt;
        t = t + dt;
        kstep += 1;
        integ = integ + tot * dt;
        // determine_procsite: the first process whose accum exceeds the
        // value (interval_search_real); past the end, the last one, walked
        // left over equal sums
        const double v = u[1] * total;
        const unsigned long long mp = __ballot(act && my_acc > v);
        int proc = mp ? __ffsll((long long)mp) - 1 : P - 1;
        if (!mp)
            while (proc > 0 && readlane_d(my_acc, proc - 1) >= readlane_d(my_acc, proc))
                --proc;
        const int np = readlane_i(n, proc);
        const double vs = u[2] * readlane_d(tot, proc);
        const double *pp = p.pre + base + (long long)proc * g.C;
        int lo_ = 0, hi_ = np;
        while (hi_ - lo_ > 64) {
            const int stride = (hi_ - lo_ + 63) >> 6;
            const int at = lo_ + (lane + 1) * stride - 1;
            const bool in = at < hi_;
            const double pv = in ? pp[at] : 0.0;
            const unsigned long long mm = __ballot(in && pv > vs);
            if (mm) {
                const int k = __ffsll((long long)mm) - 1;
                hi_ = lo_ + (k + 1) * stride - 1;
                lo_ = lo_ + k * stride;
            } else {
                lo_ = lo_ + (int)__popcll(__ballot(in)) * stride;
            }
        }
        int pos;
        {
            const int at = lo_ + lane;
            const bool in = at < hi_;
            const double pv = in ? pp[at] : 0.0;
            const unsigned long long mm = __ballot(in && pv > vs);
            pos = mm ? lo_ + __ffsll((long long)mm) - 1 : hi_;
        }
        if (pos >= np) {
            pos = np - 1;
            while (pos > 0 && pp[pos - 1] >= pp[pos])
                --pos;
        }
        const int cell = p.sites[base + (long long)proc * g.C + pos] - 1;
        int x, y, z;
        g.xyz(cell, x, y, z);
        if (lane == proc)
            pstat += 1;
        // run_proc_<proc>: the owner lane's deletes
        int dirty = 0x7fffffff;
        bool corrupt = false;
        const int own = prog[p.off_owner + proc * P + q];
        const int *w = prog + (own & 0xFFFF);
        const int nd = act ? (own >> 16) & 0xFF : 0;
        const int nrest = act ? (own >> 24) & 0xFF : 0;
        for (int k = 0; k < nd; ++k, ++w) {
            const int c = g.at(x, y, z, sx8(w[0], 2), sx8(w[0], 10), sx8(w[0], 18));
            if (brow[c]) {
                const int d = otf_del(srow, brow, rrow, n, run, c);
                if (d < 0)
                    corrupt = true;
                else
                    dirty = min(dirty, d);
            }
        }
        // lattice writes
        const int wt = prog[p.off_writes + proc];
        bool bad = false;
        if (lane < (wt >> 16)) {
            const int *ww = prog + (wt & 0xFFFF) + 2 * lane;
            const int c = g.at(x, y, z, sx8(ww[0], 0), sx8(ww[0], 8), sx8(ww[0], 16));
            const int site = p.spuck * c + ((ww[0] >> 24) & 0xFF);
            bad = slat[site] != (ww[1] & 0xFFFF);
            slat[site] = (signed char)(ww[1] >> 16);
        }
        __syncthreads();
        // the owner lane's rate updates, which come before its adds: in
        // batches of up to eight, their inverse indices read together, then
        // their old rates, then the new ones applied in program order
        int k = 0;
        for (;;) {
            int m = 0;
            while (m < 8 && k + m < nrest && (w[m] & 3) == 2)
                ++m;
            if (m == 0)
                break;
            int cx8[8], cy8[8], cz8[8], c8[8], bk8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j < m) {
                    cx8[j] = otf_wrap(x + sx8(w[j], 2), g.X);
                    cy8[j] = otf_wrap(y + sx8(w[j], 10), g.Y);
                    cz8[j] = otf_wrap(z + sx8(w[j], 18), g.Z);
                    c8[j] = g.idx(cx8[j], cy8[j], cz8[j]);
                    bk8[j] = brow[c8[j]];
                }
            }
            double old8[8], new8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m && bk8[j] > 0 && bk8[j] <= n)
                    old8[j] = rrow[bk8[j] - 1];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j < m) {
                    if (bk8[j] > n) {
                        corrupt = true;
                    } else if (bk8[j]) {
                        const int at = bk8[j] - 1;
                        double old = old8[j];
#pragma unroll
                        for (int j2 = 0; j2 < j; ++j2)       // the same cell twice in one batch
                            if (c8[j2] == c8[j] && bk8[j2] == bk8[j])
                                old = new8[j2];
                        new8[j] = otf_gr(prog, p, g, slat, tab, q, cx8[j], cy8[j], cz8[j]);
                        run = run + new8[j] - old;
                        rrow[at] = new8[j];
                        if (new8[j] != old)
                            dirty = min(dirty, at);
                    }
                }
            }
            w += m;
            k += m;
        }
        // then its adds, in program order
        for (; k < nrest; ++k) {
            const int op = w[0];
            const int cx = otf_wrap(x + sx8(op, 2), g.X), cy = otf_wrap(y + sx8(op, 10), g.Y),
                      cz = otf_wrap(z + sx8(op, 18), g.Z);
            const int c = g.idx(cx, cy, cz);
            if ((op & 3) == 2) {
                if (brow[c] > n) {
                    corrupt = true;
                } else if (brow[c]) {
                    const int at = brow[c] - 1;
                    const double rate = otf_gr(prog, p, g, slat, tab, q, cx, cy, cz);
                    const double old = rrow[at];
                    run = run + rate - old;
                    rrow[at] = rate;
                    if (rate != old)
                        dirty = min(dirty, at);
                }
                w += 1;
            } else {
                const int ncond = (op >> 26) & 0x3F;
                if (n >= g.C) {
                    corrupt = true;             // an add to a full list
                } else if (otf_conds(p, g, slat, w + 1, ncond, x, y, z)) {
                    const double rate = otf_gr(prog, p, g, slat, tab, q, cx, cy, cz);
                    srow[n] = c + 1;
                    brow[c] = n + 1;
                    rrow[n] = rate;
                    run = run + rate;
                    dirty = min(dirty, n);
                    n += 1;
                }
                w += 1 + ncond;
            }
        }
        otf_resum_rows(p.rm, p.pre, base, g.C, lane, act, dirty, n);
        if (dirty != 0x7fffffff)
            tot = n > 0 ? prow[n - 1] : 0.0;
        if (__ballot(bad)) {
            st = KMCG_REPLICA_SPECIES_MISMATCH;
            break;
        }
        if (__ballot(corrupt)) {
            st = KMCG_REPLICA_CORRUPT;
            break;
        }
    }
    __syncthreads();
    for (int i = lane; i < V; i += 64)
        glat[i] = slat[i];
    if (act) {
        p.nsites[slot] = n;
        p.procstat[slot] = pstat;
        p.integ[slot] = integ;
        p.accum[slot] = my_acc;
        p.runsum[slot] = run;
    }
    if (lane == 0) {
        p.kmc_time[r] = t;
        p.kmc_time_step[r] = t_step;
        p.kmc_step[r] = kstep;
        p.ctr[r] = ctr;
        p.status[r] = st;
    }
}

} // namespace

// ===========================================================================
// host side
// ===========================================================================
struct kmcg_otf_model {
    std::vector<int32_t> blob;
    int P = 0, S = 0, spuck = 0, K = 0, T = 0;
    int reach[3] = {0, 0, 0};        // largest |offset| per dimension
    std::vector<int> prog;           // the device program (see build)
    int off_owner = 0, off_writes = 0, off_byst = 0, off_tbase = 0, off_touch = 0;
};

struct kmcg_otf_batch {
    const kmcg_otf_model *m = nullptr;
    int n = 0, device = 0, maxV = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    bool timed = false;
    std::vector<int> dims;           // [n][4]
    std::vector<long long> lat_off, list_off;
    long long lat_total = 0, list_total = 0;
    int *d_prog = nullptr, *d_dims = nullptr, *d_rep = nullptr;
    long long *d_lat_off = nullptr, *d_list_off = nullptr;
    signed char *d_lat = nullptr;
    int *d_sites = nullptr, *d_back = nullptr, *d_nsites = nullptr, *d_status = nullptr;
    double *d_rm = nullptr, *d_pre = nullptr, *d_integ = nullptr, *d_accum = nullptr, *d_runsum = nullptr;
    double *d_table = nullptr, *d_time = nullptr, *d_dt = nullptr;
    long long *d_procstat = nullptr, *d_step = nullptr;
    unsigned long long *d_ctr = nullptr;
    unsigned *d_seed = nullptr, *d_stream = nullptr;
    std::vector<double> rates, userpar;  // per replica, as last set
};

namespace {

// x ** n for an integer n: binary powering from the low bit, 1/r for a
// negative n (the form the reference's compiler emits for real ** integer)
template <class F>
F otf_powi(F a, long long n)
{
    unsigned long long u = n < 0 ? (unsigned long long)(-n) : (unsigned long long)n;
    F r = F(1);
    while (u) {
        if (u & 1)
            r = r * a;
        u >>= 1;
        if (u)
            a = a * a;
    }
    return n < 0 ? F(1) / r : r;
}

long long otf_ipow(long long a, long long n)
{
    if (n < 0)
        return a == 1 ? 1 : (a == -1 ? ((n % 2 == 0) ? 1 : -1) : 0);
    long long r = 1;
    while (n-- > 0)
        r *= a;
    return r;
}

struct OtfVal {
    int kind;                        // 0 integer, 1 default real, 2 double
    long long i;
    float f;
    double d;
};

// one rate program (kmos_amd/otf.py opcodes) on the given counts
bool otf_eval(const int32_t *code, int ncode, const int *counts, double base, const double *userpar,
              int naux, double *out)
{
    std::vector<OtfVal> st;
    std::vector<double> aux(std::max(naux, 1), 0.0);
    int i = 0;
    while (i < ncode) {
        const int op = code[i];
        if (op == 0)
            break;
        OtfVal v{};
        switch (op) {
        case 1: v.kind = 0; v.i = code[i + 1]; st.push_back(v); i += 2; continue;
        case 2: v.kind = 1; memcpy(&v.f, &code[i + 1], 4); st.push_back(v); i += 2; continue;
        case 3: v.kind = 2; v.d = base; st.push_back(v); i += 1; continue;
        case 4: v.kind = 2; v.d = userpar[code[i + 1]]; st.push_back(v); i += 2; continue;
        case 5: v.kind = 0; v.i = counts[code[i + 1]]; st.push_back(v); i += 2; continue;
        case 6: v.kind = 2; v.d = aux[code[i + 1]]; st.push_back(v); i += 2; continue;
        case 7: if (st.empty()) return false; aux[code[i + 1]] = st.back().d; st.pop_back(); i += 2; continue;
        case 46: v.kind = 2; memcpy(&v.d, &code[i + 1], 8); st.push_back(v); i += 3; continue;
        default: break;
        }
        i += 1;
        if (st.empty())
            return false;
        OtfVal &a = st.back();
        if (op == 8) { a.f = (float)a.i; a.kind = 1; }
        else if (op == 9) { a.d = (double)a.i; a.kind = 2; }
        else if (op == 10) { a.d = (double)a.f; a.kind = 2; }
        else if (op >= 11 && op <= 13) {
            if (a.kind == 0) a.i = -a.i;
            else if (a.kind == 1) a.f = -a.f;
            else a.d = -a.d;
        } else if (op >= 20 && op <= 36) {
            if (st.size() < 2)
                return false;
            const OtfVal b = st.back();
            st.pop_back();
            OtfVal &l = st.back();
            if (op <= 31) {
                const int k = (op - 20) % 3, bop = op - k;
                if (k == 0) {
                    if (bop == 20) l.i = l.i + b.i;
                    else if (bop == 23) l.i = l.i - b.i;
                    else if (bop == 26) l.i = l.i * b.i;
                    else l.i = b.i ? l.i / b.i : 0;
                } else if (k == 1) {
                    if (bop == 20) l.f = l.f + b.f;
                    else if (bop == 23) l.f = l.f - b.f;
                    else if (bop == 26) l.f = l.f * b.f;
                    else l.f = l.f / b.f;
                } else {
                    if (bop == 20) l.d = l.d + b.d;
                    else if (bop == 23) l.d = l.d - b.d;
                    else if (bop == 26) l.d = l.d * b.d;
                    else l.d = l.d / b.d;
                }
            } else if (op == 32) l.i = otf_ipow(l.i, b.i);
            else if (op == 33) l.f = otf_powi<float>(l.f, b.i);
            else if (op == 34) l.d = otf_powi<double>(l.d, b.i);
            else if (op == 35) l.f = powf(l.f, b.f);
            else l.d = pow(l.d, b.d);
        } else if (op >= 40 && op <= 45) {
            const int dbl = (op - 40) % 2, fn = (op - 40) / 2;
            if (dbl) a.d = fn == 0 ? exp(a.d) : (fn == 1 ? log(a.d) : sqrt(a.d));
            else a.f = fn == 0 ? expf(a.f) : (fn == 1 ? logf(a.f) : sqrtf(a.f));
        } else {
            return false;
        }
    }
    if (st.empty() || st.back().kind != 2)
        return false;
    *out = st.back().d;
    return true;
}

// packed words of the device program (see the layout above)
int otf_site_word(int dx, int dy, int dz, int n1)
{
    return (dx & 0xFF) | (dy & 0xFF) << 8 | (dz & 0xFF) << 16 | (n1 - 1) << 24;
}

int otf_op_word(int kind, int dx, int dy, int dz, int ncond)
{
    return kind | (dx & 0xFF) << 2 | (dy & 0xFF) << 10 | (dz & 0xFF) << 18 | ncond << 26;
}

int otf_cond_word(int dx, int dy, int dz, int n1, int sp)
{
    return (dx & 0x7F) | (dy & 0x7F) << 7 | (dz & 0x7F) << 14 | (n1 - 1) << 21 | sp << 26;
}

int otf_build(kmcg_otf_model *m)
{
    const std::vector<int32_t> &B = m->blob;
    const int P = m->P, S = m->S;
    if (m->spuck > 32 || S > 64)
        return fail(KMCG_EUNSUPPORTED, "otf engine: at most 32 sites per cell and 64 species");
    auto in = [&](long long off, long long n) { return off >= 0 && n >= 0 && off + n <= (long long)B.size(); };
    auto small = [](int v, int lim) { return v >= -lim && v <= lim; };
    std::vector<int> &g = m->prog;
    g.assign(B.begin() + B[8], B.begin() + B[8] + m->spuck);   // site defaults first
    // flat run_proc programs -> per (exec, owner) lists and per-exec writes
    std::vector<std::vector<std::vector<int>>> own(P, std::vector<std::vector<int>>(P));
    std::vector<std::vector<int>> ndel(P, std::vector<int>(P, 0)), nrest(P, std::vector<int>(P, 0));
    std::vector<std::vector<int>> writes(P);
    auto reach = [&](int dx, int dy, int dz) {
        m->reach[0] = std::max(m->reach[0], std::abs(dx));
        m->reach[1] = std::max(m->reach[1], std::abs(dy));
        m->reach[2] = std::max(m->reach[2], std::abs(dz));
    };
    // an add op and its conditions, packed; false when out of the packing's range
    auto add_op = [&](std::vector<int> &out, const int32_t *op) {
        if (op[5] > 63 || !small(op[2], 127) || !small(op[3], 127) || !small(op[4], 127))
            return false;
        out.push_back(otf_op_word(3, op[2], op[3], op[4], op[5]));
        for (int j = 0; j < op[5]; ++j) {
            const int32_t *c = op + 6 + 5 * j;
            if (c[3] < 1 || c[3] > m->spuck || c[4] < 0 || c[4] >= S || !small(c[0], 63) ||
                !small(c[1], 63) || !small(c[2], 63))
                return false;
            reach(c[0], c[1], c[2]);
            out.push_back(otf_cond_word(c[0], c[1], c[2], c[3], c[4]));
        }
        return true;
    };
    auto parse = [&](long long off, int nops, auto &&emit) -> bool {
        long long k = off;
        for (int i = 0; i < nops; ++i) {
            if (!in(k, 1))
                return false;
            const int kind = B[k];
            const int len = kind == 1 ? 7 : (kind == 3 ? (in(k, 6) ? 6 + 5 * B[k + 5] : -1) : 5);
            if (len < 0 || !in(k, len) || kind < 0 || kind > 3)
                return false;
            if (!emit(&B[k]))
                return false;
            k += len;
        }
        return true;
    };
    for (int p = 0; p < P; ++p) {
        const long long off = B[9] + 2 * p;
        if (!in(off, 2))
            return fail(KMCG_EINVAL, "otf program table out of range");
        bool after_write = false;
        const bool ok = parse(B[off + 1], B[off], [&](const int32_t *op) {
            const int kind = op[0];
            if (kind == 1) {
                if (op[5] < 0 || op[5] >= S || op[6] < 0 || op[6] >= S || op[4] < 1 || op[4] > m->spuck ||
                    !small(op[1], 127) || !small(op[2], 127) || !small(op[3], 127))
                    return false;
                writes[p].push_back(otf_site_word(op[1], op[2], op[3], op[4]));
                writes[p].push_back(op[5] | op[6] << 16);
                reach(op[1], op[2], op[3]);
                after_write = true;
                return true;
            }
            const int q = op[1];
            if (q < 0 || q >= P || !small(op[2], 127) || !small(op[3], 127) || !small(op[4], 127))
                return false;
            reach(op[2], op[3], op[4]);
            if (kind == 0) {
                if (after_write || nrest[p][q] > 0)
                    return false;                    // deletes come first
                own[p][q].push_back(otf_op_word(0, op[2], op[3], op[4], 0));
                ndel[p][q] += 1;
            } else if (kind == 2) {
                own[p][q].push_back(otf_op_word(2, op[2], op[3], op[4], 0));
                nrest[p][q] += 1;
            } else {
                if (!add_op(own[p][q], op))
                    return false;
                nrest[p][q] += 1;
            }
            return true;
        });
        if (!ok)
            return fail(KMCG_EINVAL, "bad otf program of process " + std::to_string(p + 1));
        if (writes[p].size() / 2 > 64)
            return fail(KMCG_EUNSUPPORTED, "more than 64 lattice writes in one process");
    }
    // touchup: per process its adds
    std::vector<std::vector<int>> touch(P);
    std::vector<int> ntouch(P, 0);
    const bool tok = parse(B[12], B[18], [&](const int32_t *op) {
        if (op[0] != 3 || op[1] < 0 || op[1] >= P)
            return false;
        ntouch[op[1]] += 1;
        return add_op(touch[op[1]], op);
    });
    if (!tok)
        return fail(KMCG_EINVAL, "bad otf touchup program");
    m->off_owner = (int)g.size();
    g.resize(g.size() + P * P);
    m->off_writes = (int)g.size();
    g.resize(g.size() + P);
    m->off_byst = (int)g.size();
    g.resize(g.size() + P);
    m->off_tbase = (int)g.size();
    g.resize(g.size() + P);
    m->off_touch = (int)g.size();
    g.resize(g.size() + P);
    for (int p = 0; p < P; ++p)
        for (int q = 0; q < P; ++q) {
            if (ndel[p][q] > 255 || nrest[p][q] > 255)
                return fail(KMCG_EUNSUPPORTED, "more than 255 ops of one process in one routine");
            g[m->off_owner + p * P + q] = (int)g.size() | ndel[p][q] << 16 | nrest[p][q] << 24;
            g.insert(g.end(), own[p][q].begin(), own[p][q].end());
        }
    for (int p = 0; p < P; ++p) {
        g[m->off_writes + p] = (int)g.size() | (int)(writes[p].size() / 2) << 16;
        g.insert(g.end(), writes[p].begin(), writes[p].end());
    }
    long long T = 0;
    for (int q = 0; q < P; ++q) {
        const long long rec = B[11] + 9 * q;
        if (!in(rec, 9) || !in(B[rec + 3], (long long)B[rec + 2] * (4 + S)) || !in(B[rec + 5], B[rec + 4]) ||
            !in(B[rec + 8], B[rec + 7]) || B[rec + 0] < 1 || B[rec + 1] != T)
            return fail(KMCG_EINVAL, "bad otf rate table of process " + std::to_string(q + 1));
        long long size = 1;
        for (int v = 0; v < B[rec + 7]; ++v)
            size *= (long long)B[B[rec + 8] + v] + 1;
        if (size != B[rec + 0])
            return fail(KMCG_EINVAL, "otf rate table size does not match its counts");
        g[m->off_byst + q] = (int)g.size() | B[rec + 2] << 16;
        long long reach_idx = 0;
        for (int k = 0; k < B[rec + 2]; ++k) {
            const int32_t *b = &B[B[rec + 3] + k * (4 + S)];
            if (b[3] < 1 || b[3] > m->spuck || !small(b[0], 127) || !small(b[1], 127) || !small(b[2], 127))
                return fail(KMCG_EINVAL, "bad bystander site");
            reach(b[0], b[1], b[2]);
            g.push_back(otf_site_word(b[0], b[1], b[2], b[3]));
            int most = 0;
            for (int s2 = 0; s2 < S; ++s2) {
                if (b[4 + s2] < 0)
                    return fail(KMCG_EINVAL, "negative bystander stride");
                most = std::max(most, (int)b[4 + s2]);
                g.push_back(b[4 + s2]);
            }
            reach_idx += most;
        }
        if (reach_idx >= size)
            return fail(KMCG_EINVAL, "bystander strides reach beyond the rate table of process " +
                                     std::to_string(q + 1));
        g[m->off_tbase + q] = (int)T;
        T += B[rec + 0];
    }
    for (int q = 0; q < P; ++q) {
        g[m->off_touch + q] = (int)g.size() | ntouch[q] << 16;
        g.insert(g.end(), touch[q].begin(), touch[q].end());
    }
    if (T > (1LL << 24))
        return fail(KMCG_EUNSUPPORTED, "otf rate tables too large");
    if (g.size() >= 65536 || (long long)g.size() * 4 > OTF_LDS / 2)
        return fail(KMCG_EUNSUPPORTED, "otf program too large for the engine (" + std::to_string(g.size()) +
                                       " words)");
    m->T = (int)T;
    return KMCG_OK;
}

OtfParams otf_params(kmcg_otf_batch *b, long long steps, int mode)
{
    OtfParams p{};
    const kmcg_otf_model *m = b->m;
    p.P = m->P; p.S = m->S; p.spuck = m->spuck; p.T = m->T; p.W = (int)m->prog.size();
    p.prog = b->d_prog;
    p.off_owner = m->off_owner; p.off_writes = m->off_writes; p.off_byst = m->off_byst;
    p.off_tbase = m->off_tbase; p.off_touch = m->off_touch;
    p.dims = b->d_dims; p.lat_off = b->d_lat_off; p.list_off = b->d_list_off;
    p.lat = b->d_lat; p.sites = b->d_sites; p.back = b->d_back; p.rm = b->d_rm; p.pre = b->d_pre;
    p.nsites = b->d_nsites; p.procstat = b->d_procstat; p.integ = b->d_integ; p.accum = b->d_accum;
    p.runsum = b->d_runsum; p.table = b->d_table; p.kmc_time = b->d_time; p.kmc_time_step = b->d_dt;
    p.kmc_step = b->d_step;
    p.ctr = b->d_ctr; p.seed = b->d_seed; p.stream = b->d_stream; p.status = b->d_status;
    p.rep_ids = b->d_rep; p.steps = steps; p.mode = mode;
    return p;
}

int otf_sync(kmcg_otf_batch *b)
{
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timed) {
        HIPCHK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
        b->timed = false;
    }
    return KMCG_OK;
}

int otf_range(const kmcg_otf_batch *b, int32_t r0, int32_t n)
{
    if (!b || r0 < 0 || n < 1 || (long long)r0 + n > b->n)
        return fail(KMCG_EINVAL, "replica range out of the batch");
    return KMCG_OK;
}

int otf_launch(kmcg_otf_batch *b, int32_t r0, int32_t n, long long steps, int mode, bool timed)
{
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i)
        ids[i] = r0 + i;
    HIPCHK(hipSetDevice(b->device));
    int rc = otf_sync(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(b->d_rep, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    OtfParams p = otf_params(b, steps, mode);
    // the steps kernel stages the program (and, when it fits, the rate
    // table) in LDS ahead of the lattice
    const size_t prog_bytes = (size_t)((p.W + 1) & ~1) * 4;
    const size_t tab_bytes = (size_t)b->m->T * 8;
    p.tab_lds = prog_bytes + tab_bytes + (size_t)b->maxV <= (size_t)OTF_LDS;
    if (timed)
        HIPCHK(hipEventRecord(b->ev0, b->stream));
    if (mode < 0 && p.tab_lds)
        hipLaunchKernelGGL(otf_steps_kernel<true>, dim3(n), dim3(64), prog_bytes + tab_bytes + b->maxV,
                           b->stream, p);
    else if (mode < 0)
        hipLaunchKernelGGL(otf_steps_kernel<false>, dim3(n), dim3(64), prog_bytes + b->maxV, b->stream, p);
    else
        hipLaunchKernelGGL(otf_setup_kernel, dim3(n), dim3(64), (size_t)b->maxV, b->stream, p);
    HIPCHK(hipGetLastError());
    if (timed) {
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        b->timed = true;
    }
    return KMCG_OK;
}

// the host evaluates each replica's tables (one per process, every count
// combination) and sends them to the device
int otf_tables(kmcg_otf_batch *b, int32_t r0, int32_t n)
{
    const kmcg_otf_model *m = b->m;
    const std::vector<int32_t> &B = m->blob;
    std::vector<double> tab((size_t)n * m->T);
    std::vector<int> counts(64);
    for (int i = 0; i < n; ++i) {
        const int r = r0 + i;
        for (int q = 0; q < m->P; ++q) {
            const long long rec = B[11] + 9 * q;
            const int size = B[rec], tbase = B[rec + 1], nused = B[rec + 7];
            if (nused > 64)
                return fail(KMCG_EUNSUPPORTED, "more than 64 bystander counts in one rate");
            for (int t = 0; t < size; ++t) {
                int rem = t;
                for (int v = 0; v < nused; ++v) {
                    counts[v] = rem % (B[B[rec + 8] + v] + 1);
                    rem /= B[B[rec + 8] + v] + 1;
                }
                double val;
                if (!otf_eval(&B[B[rec + 5]], B[rec + 4], counts.data(), b->rates[(size_t)r * m->P + q],
                              b->userpar.data() + (size_t)r * m->K, B[rec + 6], &val))
                    return fail(KMCG_EINVAL, "bad rate program of process " + std::to_string(q + 1));
                if (!(val >= 0.0))
                    return fail(KMCG_EINVAL, "otf rate of process " + std::to_string(q + 1) +
                                             " is negative or not a number");
                tab[(size_t)i * m->T + tbase + t] = val;
            }
        }
    }
    HIPCHK(hipSetDevice(b->device));
    int rc = otf_sync(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(b->d_table + (size_t)r0 * m->T, tab.data(), tab.size() * sizeof(double),
                          hipMemcpyHostToDevice, b->stream));
    return otf_sync(b);
}

template <class T>
int otf_copy_out(kmcg_otf_batch *b, T *dst, const T *src, size_t count)
{
    int rc = otf_sync(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
    return KMCG_OK;
}

} // namespace

extern "C" {

int kmcg_otf_model_load(const int32_t *blob, int64_t nwords, kmcg_otf_model **out)
{
    if (!blob || !out || nwords < 32 || nwords > (1LL << 28))
        return fail(KMCG_EINVAL, "bad otf model blob");
    if ((uint32_t)blob[0] != 0x4B4D4347u || blob[1] != 1 || blob[2] != 2 || blob[16] != nwords)
        return fail(KMCG_EINVAL, "not an otf model blob (magic, version, backend or size)");
    auto m = new kmcg_otf_model();
    m->blob.assign(blob, blob + nwords);
    m->P = blob[3];
    m->S = blob[4];
    m->spuck = blob[6];
    m->K = blob[10];
    if (m->P < 1 || m->P > OTF_MAX_P || m->S < 1 || m->S > 64 || m->spuck < 1 || m->K < 0 ||
        blob[5] != -1) {
        delete m;
        return fail(KMCG_EUNSUPPORTED, "otf engine: 1..64 processes, one layer");
    }
    const int rc = otf_build(m);
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return KMCG_OK;
}

void kmcg_otf_model_free(kmcg_otf_model *m) { delete m; }

void kmcg_otf_batch_free(kmcg_otf_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    void *ptrs[] = {b->d_prog, b->d_dims, b->d_rep, b->d_lat_off, b->d_list_off, b->d_lat, b->d_sites,
                    b->d_back, b->d_nsites, b->d_status, b->d_rm, b->d_pre, b->d_integ, b->d_accum,
                    b->d_runsum, b->d_table, b->d_time, b->d_dt, b->d_procstat, b->d_step, b->d_ctr,
                    b->d_seed, b->d_stream};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

int kmcg_otf_batch_create(const kmcg_otf_model *m, int32_t n, const int32_t *sizes, int32_t device,
                          kmcg_otf_batch **out)
{
    if (!m || n < 1 || !sizes || !out)
        return fail(KMCG_EINVAL, "bad otf batch arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return fail(KMCG_EINVAL, "no GPU device visible");
    if (device < 0 || device >= count)
        return fail(KMCG_EINVAL, "device index out of range");
    auto b = new kmcg_otf_batch();
    b->m = m;
    b->n = n;
    b->device = device;
    b->dims.resize(4 * (size_t)n);
    b->lat_off.resize(n);
    b->list_off.resize(n);
    for (int r = 0; r < n; ++r) {
        const int X = sizes[3 * r], Y = sizes[3 * r + 1], Z = sizes[3 * r + 2];
        const long long room = OTF_LDS - (long long)((m->prog.size() + 1) & ~(size_t)1) * 4;
        if (X < 1 || Y < 1 || Z < 1 || X < m->reach[0] || Y < m->reach[1] || Z < m->reach[2] ||
            (long long)X * Y * Z * m->spuck > std::min<long long>(OTF_MAX_V, room)) {
            delete b;
            return fail(KMCG_EINVAL, "lattice size of replica " + std::to_string(r) +
                                     " outside the engine's range (each extent at least the model's "
                                     "largest offset; at most " + std::to_string(std::min<long long>(OTF_MAX_V, room)) +
                                     " sites beside this model's program in 64 KiB of LDS)");
        }
        const int C = X * Y * Z;
        b->dims[4 * r] = X; b->dims[4 * r + 1] = Y; b->dims[4 * r + 2] = Z; b->dims[4 * r + 3] = C;
        b->lat_off[r] = b->lat_total;
        b->lat_total += (long long)C * m->spuck;
        b->list_off[r] = b->list_total;
        b->list_total += (long long)C * m->P;
        b->maxV = std::max(b->maxV, C * m->spuck);
    }
    b->rates.assign((size_t)n * m->P, 0.0);
    b->userpar.assign((size_t)n * std::max(m->K, 1), 0.0);
    auto bail = [&](int code, const std::string &msg) {
        kmcg_otf_batch_free(b);
        return fail(code, msg);
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return bail(KMCG_EHIP, "stream or event creation failed");
    const size_t P = m->P;
    struct A { void **p; size_t bytes; } allocs[] = {
        {(void **)&b->d_prog, m->prog.size() * sizeof(int)},
        {(void **)&b->d_dims, b->dims.size() * sizeof(int)},
        {(void **)&b->d_rep, (size_t)n * sizeof(int)},
        {(void **)&b->d_lat_off, (size_t)n * sizeof(long long)},
        {(void **)&b->d_list_off, (size_t)n * sizeof(long long)},
        {(void **)&b->d_lat, (size_t)b->lat_total},
        {(void **)&b->d_sites, (size_t)b->list_total * sizeof(int)},
        {(void **)&b->d_back, (size_t)b->list_total * sizeof(int)},
        {(void **)&b->d_rm, (size_t)b->list_total * sizeof(double)},
        {(void **)&b->d_pre, (size_t)b->list_total * sizeof(double)},
        {(void **)&b->d_nsites, n * P * sizeof(int)},
        {(void **)&b->d_status, (size_t)n * sizeof(int)},
        {(void **)&b->d_integ, n * P * sizeof(double)},
        {(void **)&b->d_accum, n * P * sizeof(double)},
        {(void **)&b->d_runsum, n * P * sizeof(double)},
        {(void **)&b->d_table, (size_t)n * std::max(m->T, 1) * sizeof(double)},
        {(void **)&b->d_time, (size_t)n * sizeof(double)},
        {(void **)&b->d_dt, (size_t)n * sizeof(double)},
        {(void **)&b->d_procstat, n * P * sizeof(long long)},
        {(void **)&b->d_step, (size_t)n * sizeof(long long)},
        {(void **)&b->d_ctr, (size_t)n * sizeof(unsigned long long)},
        {(void **)&b->d_seed, (size_t)n * sizeof(unsigned)},
        {(void **)&b->d_stream, (size_t)n * sizeof(unsigned)},
    };
    for (auto &a : allocs)
        if (hipMalloc(a.p, std::max(a.bytes, (size_t)8)) != hipSuccess)
            return bail(KMCG_ENOMEM, "device allocation failed for an otf batch");
    if (hipMemcpy(b->d_prog, m->prog.data(), m->prog.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_dims, b->dims.data(), b->dims.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_lat_off, b->lat_off.data(), n * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_list_off, b->list_off.data(), n * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_nsites, 0, n * P * sizeof(int)) != hipSuccess ||
        hipMemset(b->d_runsum, 0, n * P * sizeof(double)) != hipSuccess ||
        hipMemset(b->d_table, 0, (size_t)n * std::max(m->T, 1) * sizeof(double)) != hipSuccess ||
        hipMemset(b->d_status, 0xff, (size_t)n * sizeof(int)) != hipSuccess)
        return bail(KMCG_EHIP, "otf batch set-up copies failed");
    *out = b;
    return KMCG_OK;
}

int kmcg_otf_init(kmcg_otf_batch *b, int32_t r0, int32_t n, const uint32_t *seeds, const uint32_t *streams)
{
    int rc = otf_range(b, r0, n);
    if (rc)
        return rc;
    if (!seeds || !streams)
        return fail(KMCG_EINVAL, "seeds and streams are required");
    if ((rc = otf_sync(b)))
        return rc;
    HIPCHK(hipMemcpy(b->d_seed + r0, seeds, n * sizeof(unsigned), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_stream + r0, streams, n * sizeof(unsigned), hipMemcpyHostToDevice));
    if ((rc = otf_launch(b, r0, n, 0, 0, false)))
        return rc;
    return otf_sync(b);
}

int kmcg_otf_set_params(kmcg_otf_batch *b, int32_t r0, int32_t n, const double *rates, const double *userpar)
{
    int rc = otf_range(b, r0, n);
    if (rc)
        return rc;
    const int P = b->m->P, K = b->m->K;
    if (!rates || (K > 0 && !userpar))
        return fail(KMCG_EINVAL, "rates (and parameters) are required");
    for (long long i = 0; i < (long long)n * P; ++i)
        if (!(rates[i] >= 0.0) || !std::isfinite(rates[i]))
            return fail(KMCG_EINVAL, "rate constants must be finite and non-negative");
    for (long long i = 0; i < (long long)n * K; ++i)
        if (!std::isfinite(userpar[i]))
            return fail(KMCG_EINVAL, "parameters must be finite");
    std::vector<double> old_r(b->rates), old_u(b->userpar);
    std::copy(rates, rates + (size_t)n * P, b->rates.begin() + (size_t)r0 * P);
    if (K > 0)
        std::copy(userpar, userpar + (size_t)n * K, b->userpar.begin() + (size_t)r0 * K);
    if ((rc = otf_tables(b, r0, n))) {
        b->rates.swap(old_r);
        b->userpar.swap(old_u);
        return rc;
    }
    int status0 = 0;
    HIPCHK(hipMemcpy(&status0, b->d_status + r0, sizeof(int), hipMemcpyDeviceToHost));
    if (status0 == -1)
        return KMCG_OK;                   // not initialised yet: init builds the lists
    if ((rc = otf_launch(b, r0, n, 0, 2, false)))
        return rc;
    return otf_sync(b);
}

int kmcg_otf_set_config(kmcg_otf_batch *b, int32_t r0, int32_t n, const int8_t *config)
{
    int rc = otf_range(b, r0, n);
    if (rc)
        return rc;
    if (!config)
        return fail(KMCG_EINVAL, "config is required");
    long long total = 0;
    for (int i = 0; i < n; ++i)
        total += (long long)b->dims[4 * (r0 + i) + 3] * b->m->spuck;
    for (long long i = 0; i < total; ++i)
        if (config[i] < 0 || config[i] >= b->m->S)
            return fail(KMCG_EINVAL, "species ids must be 0..S-1");
    std::vector<int> st(n);
    if ((rc = otf_copy_out(b, st.data(), b->d_status + r0, n)))
        return rc;
    for (int i = 0; i < n; ++i)
        if (st[i] == -1)
            return fail(KMCG_EINVAL, "set_config before init");
    HIPCHK(hipMemcpy(b->d_lat + b->lat_off[r0], config, total, hipMemcpyHostToDevice));
    if ((rc = otf_launch(b, r0, n, 0, 1, false)))
        return rc;
    return otf_sync(b);
}

int kmcg_otf_run(kmcg_otf_batch *b, int32_t r0, int32_t n, int64_t steps)
{
    int rc = otf_range(b, r0, n);
    if (rc)
        return rc;
    if (steps < 0)
        return fail(KMCG_EINVAL, "steps must be non-negative");
    std::vector<int> st(n);
    if ((rc = otf_copy_out(b, st.data(), b->d_status + r0, n)))
        return rc;
    for (int i = 0; i < n; ++i)
        if (st[i] == -1)
            return fail(KMCG_EINVAL, "run before init");
    return otf_launch(b, r0, n, steps, -1, true);
}

int kmcg_otf_sync(kmcg_otf_batch *b)
{
    if (!b)
        return fail(KMCG_EINVAL, "null batch");
    return otf_sync(b);
}

int kmcg_otf_last_run_ms(kmcg_otf_batch *b, float *ms)
{
    if (!b || !ms)
        return fail(KMCG_EINVAL, "bad arguments");
    int rc = otf_sync(b);
    *ms = b->last_ms;
    return rc;
}

int kmcg_otf_get_scalars(kmcg_otf_batch *b, int32_t r, double *kmc_time, double *kmc_time_step, int64_t *kmc_step,
                         int32_t *status)
{
    int rc = otf_range(b, r, 1);
    if (rc)
        return rc;
    if ((rc = otf_copy_out(b, kmc_time, b->d_time + r, 1)) || (rc = otf_copy_out(b, kmc_time_step, b->d_dt + r, 1)))
        return rc;
    long long s = 0;
    if ((rc = otf_copy_out(b, &s, b->d_step + r, 1)))
        return rc;
    *kmc_step = s;
    return otf_copy_out(b, (int *)status, b->d_status + r, 1);
}

int kmcg_otf_get_lattice(kmcg_otf_batch *b, int32_t r, int8_t *out)
{
    int rc = otf_range(b, r, 1);
    if (rc)
        return rc;
    return otf_copy_out(b, (signed char *)out, b->d_lat + b->lat_off[r], (size_t)b->dims[4 * r + 3] * b->m->spuck);
}

int kmcg_otf_get_counters(kmcg_otf_batch *b, int32_t r, int32_t *nr_of_sites, int64_t *procstat, double *integ,
                          double *accum)
{
    int rc = otf_range(b, r, 1);
    if (rc)
        return rc;
    const size_t P = b->m->P;
    if ((rc = otf_copy_out(b, (int *)nr_of_sites, b->d_nsites + r * P, P)))
        return rc;
    if ((rc = otf_copy_out(b, (long long *)procstat, b->d_procstat + r * P, P)))
        return rc;
    if ((rc = otf_copy_out(b, integ, b->d_integ + r * P, P)))
        return rc;
    return otf_copy_out(b, accum, b->d_accum + r * P, P);
}

int kmcg_otf_get_avail(kmcg_otf_batch *b, int32_t r, int32_t *sites, int32_t *back, double *rates_matrix)
{
    int rc = otf_range(b, r, 1);
    if (rc)
        return rc;
    const kmcg_otf_model *m = b->m;
    const int P = m->P, C = b->dims[4 * r + 3], V = C * m->spuck;
    std::vector<int> s((size_t)P * C), bk((size_t)P * C);
    std::vector<double> rm((size_t)P * C), run(P);
    if ((rc = otf_copy_out(b, s.data(), b->d_sites + b->list_off[r], s.size())) ||
        (rc = otf_copy_out(b, bk.data(), b->d_back + b->list_off[r], bk.size())) ||
        (rc = otf_copy_out(b, rm.data(), b->d_rm + b->list_off[r], rm.size())) ||
        (rc = otf_copy_out(b, run.data(), b->d_runsum + (size_t)r * P, (size_t)P)))
        return rc;
    // avail_sites(p, i, 1): site numbers (site 1 of the cell), avail_sites(p,
    // site, 2): list positions; rates_matrix(p, 1:V) and (p, V+1)
    for (int p = 0; p < P; ++p) {
        for (int i = 0; i < V; ++i) {
            sites[(size_t)p * V + i] = 0;
            back[(size_t)p * V + i] = 0;
            rates_matrix[(size_t)p * (V + 1) + i] = 0.0;
        }
        for (int i = 0; i < C; ++i) {
            const int cell1 = s[(size_t)p * C + i];
            sites[(size_t)p * V + i] = cell1 ? (cell1 - 1) * m->spuck + 1 : 0;
            back[(size_t)p * V + (size_t)i * m->spuck] = bk[(size_t)p * C + i];
            rates_matrix[(size_t)p * (V + 1) + i] = rm[(size_t)p * C + i];
        }
        rates_matrix[(size_t)p * (V + 1) + V] = run[p];
    }
    return KMCG_OK;
}

int kmcg_otf_get_table(kmcg_otf_batch *b, int32_t r, double *out)
{
    int rc = otf_range(b, r, 1);
    if (rc)
        return rc;
    return otf_copy_out(b, out, b->d_table + (size_t)r * b->m->T, (size_t)b->m->T);
}

int kmcg_otf_model_info(const kmcg_otf_model *m, int32_t *info4)
{
    if (!m || !info4)
        return fail(KMCG_EINVAL, "bad arguments");
    info4[0] = m->P;
    info4[1] = m->S;
    info4[2] = m->K;
    info4[3] = m->T;
    return KMCG_OK;
}

} // extern "C"
