// kmcg_engine.hip -- MI355X kMC engine: kernels and the C ABI of kmcg.h.
//
// Exact serial-trajectory kMC (kmos local_smart / lat_int semantics), one
// wavefront per replica, thousands of replicas per launch.  See
// kmcg_device.h for the data layout and the lane-ownership scheme.
//
// Reference semantics followed (paths relative to the kmos tree):
//   step loop            kmos/fortran_src/proclist_generic_subroutines.mpy:1-80
//   accum/integ rates    kmos/fortran_src/base.mpy:603-645
//   process/site choice  kmos/fortran_src/base.mpy:1075-1120, 1234-1338
//   clock                kmos/fortran_src/base.mpy:1123-1161
//   add/del_proc         kmos/fortran_src/base.mpy:211-301 (lat_int: base_lat_int.mpy:197-264)
//   put/take/touchup     kmos/io/__init__.py:2219-2443 (compiled to segments)
//   lat_int groups       kmos/io/__init__.py:1732-2057 (compiled to segments)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "kmcg.h"
#include "kmcg_device.h"
#include "kmcg_otf.h"

using namespace kmcg;

// ===========================================================================
// device code
// ===========================================================================
namespace {

// Phase profile of the step loop (a build with -DKMCG_PHASES=1 only): each
// wave adds the shader-clock ticks between marks to one of five counters
// (0 select, 1 site read, 2 post, 3 replay, 4 the rest of the step) and
// counts the replay's blocks (batches of HBM reads) and replays.  The
// clock reads wait for outstanding scalar and LDS work, so the instrumented
// loop runs a little slower than the plain one.
#ifndef KMCG_PHASES
#define KMCG_PHASES 0
#endif
#if KMCG_PHASES
#define PH_MARK(W, i)                                                  \
    do {                                                               \
        const unsigned long long t_ = __builtin_amdgcn_s_memtime();     \
        (W).tph[i] += t_ - (W).tmark;                                  \
        (W).tmark = t_;                                                \
    } while (0)
#else
#define PH_MARK(W, i) \
    do {              \
    } while (0)
#endif

constexpr int MODE_INIT = 0;       // touchups in k, j, i order (initialize_state)
constexpr int MODE_ADJUST = 1;     // touchups with x outermost (_adjust_database)
constexpr int MODE_TOUCHUP = 2;    // one touchup at (arg0..arg3)
constexpr int MODE_RUN_PROC = 3;   // run_proc_nr(arg0, arg1)
constexpr int MODE_NEXT_STEP = 4;  // get_next_kmc_step -> out[0..1]

// One replica driven by one wavefront.
template <int PPL, int BITS>
struct Wave {
    const int *prog;   // packed op program (LDS)
    Dims d;
    Magic mx, my, ms;  // division by X, Y, spuck
    Lattice<BITS> lat;
    int *mbox;         // queues [mrows][64*PPL]: entry j of process p at j*64*PPL + p
    int *cnt;          // [64*PPL] queue lengths
    int *win;          // [64][WIN] list-tail window of each lane (replay)
    int *occ;          // [S][spuck] species counts, kept while `track`
    int *bvq;          // [RBLK][64*PPL] read-ahead back-indices (LIP form)
    bool track;
    bool chk;          // lists match the lattice: disable checks may test conditions
    bool local_smart;
    int *avail;        // sites[P][V+1] then back[P][V+1]
    long long V1;
    int P, S, lane, mstride, mrows;
    int off_seg, off_proc, off_touch, off_nli, off_nlitab;
    int nr[PPL];
    double rate[PPL];
    double integ[PPL];
    long long pstat[PPL];
#if KMCG_PHASES
    unsigned long long tph[5];
    unsigned long long nblk, napply;   // replay blocks, replays
    unsigned long long tmark;
#endif

    __device__ __forceinline__ int *sites(int p) const { return avail + (long long)p * V1; }
    __device__ __forceinline__ int *back(int p) const { return avail + (long long)(P + p) * V1; }
};

template <int PPL, class T>
__device__ __forceinline__ T slot_get(const T (&a)[PPL], int k)
{
    T v = a[0];
#pragma unroll
    for (int i = 1; i < PPL; ++i)
        if (k == i) v = a[i];
    return v;
}

// nli_<group>(cell): walk the case tree (kmos/io/__init__.py:1985-2057).
template <int PPL, int BITS, bool IN>
__device__ int nli_eval(const Wave<PPL, BITS> &W, int group, int cx, int cy, int cz)
{
    const int *prog = W.prog;
    const int lc = W.d.lin(cx, cy, cz);
    const int th = prog[W.off_nlitab + group];
    if (th >= 0) {
        // tabulated: read the K coordinates, their species, one table byte
        const int hdr = prog[th], K = hdr & 0xFF, S = (hdr >> 8) & 0xFF;
        int cw[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            cw[k] = prog[th + 1 + k];
        int sp[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = (k < K) ? cw[k] : (1 << 24);
            sp[k] = W.lat.get(site_from<IN>(W.d, lc, cx, cy, cz, unpack_x(c), unpack_y(c), unpack_z(c),
                                            (c >> 24) & 0xFF));
        }
        int idx = 0;
#pragma unroll
        for (int k = 7; k >= 0; --k)
            idx = (k < K) ? idx * S + sp[k] : idx;
        const unsigned char *tab = (const unsigned char *)(prog + th + 1 + K);
        return tab[idx];
    }
    int pos = prog[W.off_nli + group];
    for (;;) {
        const int *nd = prog + pos;
        const int nc = nd[1];
        if (nc == -1)
            return nd[2];
        const int c = nd[0];
        const int sp = W.lat.get(site_from<IN>(W.d, lc, cx, cy, cz, unpack_x(c), unpack_y(c), unpack_z(c),
                                               (c >> 24) & 0xFF));
        int next = -1;
        for (int k = 0; k < nc; ++k) {
            const int *e = nd + 2 + 3 * k;
            if (sp >= 0 && ((e[0] >> sp) & 1)) {
                if (e[1] == 1)
                    return e[2];
                next = e[2];
                break;
            }
        }
        if (next < 0)
            return 0;
        pos = next;
    }
}

// ---------------------------------------------------------------------------
// Replay of a block of queued entries of one process (add_proc/del_proc,
// base.mpy:211-301; lat_int: base_lat_int.mpy:197-264).
//
// v[0..m) are the entries in reference order: deletions and additions may
// interleave.  One batch of HBM reads serves the block: back[site] of every
// site to delete and the last D list slots (D = number of deletions).  The
// swap-with-last of del_proc is replayed in registers over a window of list
// slots [nr0-RBLK+1, nr0+RBLK] (the only slots a later entry of the block
// can read), tracking the slots and back-indices the block itself changes,
// and every change is written through to HBM in program order.
// ---------------------------------------------------------------------------
constexpr int RBLK = 4;
constexpr int WIN = 2 * RBLK;

// 32-bit byte offsets into one replica's avail tables, from the uniform
// table base: one VALU instruction per address
struct Rows {
    char *base;        // the replica's avail tables (uniform)
    unsigned so, bo;   // byte offsets of this lane's sites[p] and back[p]
    __device__ __forceinline__ int &s(int i) const { return *(int *)(base + (so + ((unsigned)i << 2))); }
    __device__ __forceinline__ int &b(int i) const { return *(int *)(base + (bo + ((unsigned)i << 2))); }
};

// w: this lane's window in LDS, slot q at w[q - (nr0 - RB + 1)].  RB (1, 2
// or RBLK) bounds the entries any lane replays in this block: the reads,
// the window and the pairwise tests are set up for RB entries only.
// List words of a lane's first replay block read ahead of the replay:
// back[site] of each entry that deletes (read by the op lane that queued
// it, qv, and stored at qi of W.bvq when the replay starts, so that the
// read is waited for only then) and the list slots nr, nr-1, ... (read
// before the post).  Nothing in the step changes them before the block
// runs.
struct RePre {
    bool on = false;
    int qv = 0, qi = -1;
    int bv[RBLK];
    int tv[RBLK];
};

// Work to run while a replay block's reads are on their way (wave-uniform,
// every lane active); the default does nothing.
struct NoMid {
    static constexpr bool active = false;
    bool armed;
    template <class Wv>
    __device__ __forceinline__ void operator()(const Wv &, int) {}
};

template <int PPL, int BITS, int RB, class Mid>
__device__ __forceinline__ void replay_block(Wave<PPL, BITS> &W, int &nr, const Rows &L, int *w,
                                             const int (&v)[RBLK], int m, int &err, Mid &mid, const RePre &pre)
{
    const int nr0 = nr;
    // per entry: deletion?, site; pairwise equal sites, computed once
    bool isdel[RB];
    int st[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        isdel[i] = i < m && (v[i] & 3) != MB_ADD;
        st[i] = v[i] >> 2;
    }
    // all reads are issued unconditionally (index 0 = the unused slot 0
    // when not needed) so that they form one batch with a single wait
    int bv[RB];
    int nd = 0;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        bv[i] = pre.on ? pre.bv[i] : L.b(isdel[i] ? st[i] : 0);
        nd += isdel[i] ? 1 : 0;
    }
    // (read ahead, every slot is read: the window's slots below nr0-nd+1
    // are written but never read)
    int tv[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j)
        tv[j] = pre.on ? pre.tv[j] : L.s((j < nd && nr0 - j >= 1) ? nr0 - j : 0);
    if (Mid::active && mid.armed) {
        // the count this lane's block leaves when every deletion finds its
        // site registered
        int dn = 0;
#pragma unroll
        for (int i = 0; i < RB; ++i)
            dn += (i < m) ? (isdel[i] ? -1 : 1) : 0;
        mid(W, nr0 + dn);
    }
    bool same[RB][RB];     // same[i][j], i < j: entry j deletes entry i's site
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = i + 1; j < RB; ++j)
            same[i][j] = isdel[j] && st[j] == st[i];
#pragma unroll
    for (int i = 0; i < RB; ++i)
        bv[i] = isdel[i] ? bv[i] : 0;
    // slots above nr0 are only read after an addition of this block wrote
    // them, slots below nr0-nd+1 are never read (their entries here are
    // written anyway, so that no read above is conditional)
#pragma unroll
    for (int j = 0; j < RB; ++j)
        w[RB - 1 - j] = tv[j];
    int cur = nr0;
    const int base = nr0 - (RB - 1);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (!__ballot(i < m))
            break;
        if (i < m) {
            const int site = st[i];
            if (!isdel[i]) {
                if (cur + 1 > W.d.V) {      // only a double add can get here
                    err |= 2;
                } else {
                    cur += 1;
                    L.s(cur) = site;
                    L.b(site) = cur;
                    w[cur - base] = site;
#pragma unroll
                    for (int j = i + 1; j < RB; ++j)
                        bv[j] = same[i][j] ? cur : bv[j];
                }
            } else {
                const int mpos = bv[i];
                if (mpos == 0) {            // not registered: can_do false
                    if ((v[i] & 3) == MB_DEL_STRICT)
                        err |= 1;
                } else {
                    const int last = cur;
                    if (mpos < last) {
                        const int moved = w[last - base];
                        L.s(mpos) = moved;
                        L.b(moved) = mpos;
                        if (mpos >= base)
                            w[mpos - base] = moved;
                        // bv of additions is unused, so no deletion test here
#pragma unroll
                        for (int j = i + 1; j < RB; ++j)
                            bv[j] = (st[j] == moved) ? mpos : bv[j];
                    }
                    L.s(last) = 0;
                    L.b(site) = 0;
#pragma unroll
                    for (int j = i + 1; j < RB; ++j)
                        bv[j] = same[i][j] ? 0 : bv[j];
                    cur = last - 1;
                }
            }
        }
    }
    nr = cur;
}

// Apply the queued entries of the step: each owner lane replays its
// processes' queues in order, RBLK entries per batch of HBM reads, and
// empties them.  With pre.on (one process slot per lane) the first block
// takes its list words from pre.tv and W.bvq instead of reading them.
template <int PPL, int BITS, bool RB2 = false, class Mid = NoMid>
__device__ __forceinline__ int apply_queues(Wave<PPL, BITS> &W, Mid &mid, RePre &pre)
{
    int err = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = W.lane + 64 * k;
        const bool own = p < W.P;
        const int *q = W.mbox + p;
        // the count and the first block are read together
        int n = W.cnt[p];
        int v[RBLK];
#pragma unroll
        for (int i = 0; i < RBLK; ++i)
            v[i] = q[min(i, W.mrows - 1) * W.mstride];
        n = own ? n : 0;
        if (PPL == 1 && pre.on) {
            if (pre.qi >= 0)
                W.bvq[pre.qi] = pre.qv;
            wave_sync();
#pragma unroll
            for (int i = 0; i < RBLK; ++i)
                pre.bv[i] = W.bvq[i * W.mstride + p];
        }
        W.cnt[p] = 0;
#if KMCG_PHASES
        W.napply += 1;
#endif
        for (int i0 = 0; __ballot(i0 < n); i0 += RBLK) {
            const int m = min(n - i0, RBLK);
#if KMCG_PHASES
            W.nblk += 1;
#endif
            if (i0 > 0) {
#pragma unroll
                for (int i = 0; i < RBLK; ++i)
                    v[i] = q[min(i0 + i, W.mrows - 1) * W.mstride];
                pre.on = false;
            }
#pragma unroll
            for (int i = 0; i < RBLK; ++i)
                v[i] = (i < m) ? v[i] : 0;
            // with an active hook (one process slot per lane) every lane
            // takes part in the block; a lane without entries reads the
            // table's first word only
            mid.armed = Mid::active && i0 == 0 && !__ballot(n > RBLK);
            if (Mid::active || m > 0) {
                const bool any = m > 0;
                int nrk = W.nr[k];
                Rows L;
                L.base = (char *)W.avail;
                L.so = any ? (unsigned)((long long)p * W.V1) << 2 : 0u;
                L.bo = any ? (unsigned)((long long)(W.P + p) * W.V1) << 2 : 0u;
                // blocks of one entry per lane are the common case; a
                // two-entry form as well (RB2) costs code size, so only the
                // plain step loop takes it
                if (__ballot(m > 2) || (!RB2 && __ballot(m > 1)))
                    replay_block<PPL, BITS, RBLK>(W, nrk, L, W.win + WIN * W.lane, v, max(m, 0), err, mid, pre);
                else if (RB2 && __ballot(m > 1))
                    replay_block<PPL, BITS, 2>(W, nrk, L, W.win + WIN * W.lane, v, max(m, 0), err, mid, pre);
                else
                    replay_block<PPL, BITS, 1>(W, nrk, L, W.win + WIN * W.lane, v, max(m, 0), err, mid, pre);
                W.nr[k] = nrk;
            }
        }
    }
    wave_sync();
    if (__ballot(err & 2))
        return ST_CORRUPT;
    return __ballot(err & 1) ? ST_DEL_ABSENT : ST_OK;
}

template <int PPL, int BITS, bool IN>
__device__ __forceinline__ int cond_species(const Wave<PPL, BITS> &W, int cw, int lb, int bx, int by, int bz)
{
    return W.lat.get(site_from<IN>(W.d, lb, bx, by, bz, (cw & 63) - 32, ((cw >> 6) & 63) - 32,
                                   ((cw >> 12) & 63) - 32, (cw >> 18) & 63));
}

// Evaluate op `op` (its first four condition words in cw) at base cell
// (bx, by, bz) on the current lattice and append it to its owner's queue.
// Ops on one process sit in different rounds, so issuing the rounds in
// order keeps the reference order within every queue.
template <int PPL, int BITS, bool IN>
__device__ __forceinline__ void eval_post(Wave<PPL, BITS> &W, bool valid, int2 op, const int (&cw)[4],
                                          const int *cd, bool as_add, int nrounds, int lb, int bx, int by, int bz,
                                          int sw, int oldw)
{
    int proc = 0, code = 0, s = 0, rnd = 0;
    if (valid) {
        int kind = op.x & 7;
        if (as_add && kind == OP_DEL_NLI)
            kind = OP_ADD_NLI;
        rnd = (op.x >> 3) & 0xFF;
        const int n = (op.x >> 11) & 0x3F;
        const int arg = (int)((unsigned)op.x >> 23);
        const int dx = sext7(op.y & 0x7F), dy = sext7((op.y >> 7) & 0x7F), dz = sext7((op.y >> 14) & 0x7F);
        if (kind == OP_DEL_IF || kind == OP_ADD_IF) {
            s = site_from<IN>(W.d, lb, bx, by, bz, dx, dy, dz, n);
            const bool del = kind == OP_DEL_IF;
            const int nc = (op.x >> 17) & 0x3F;
            bool ok = true;
            // an addition tests its conditions on the lattice after the
            // segment's write; a disable check (while the lists match the
            // lattice) tests its process's conditions on the lattice before
            // it: the process is registered exactly when they hold
            if (!del || (W.chk && nc > 0)) {
                int cs[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int w = (c < nc) ? cw[c] : (32 | 32 << 6 | 32 << 12 | 1 << 18);
                    cs[c] = site_from<IN>(W.d, lb, bx, by, bz, (w & 63) - 32, ((w >> 6) & 63) - 32,
                                          ((w >> 12) & 63) - 32, (w >> 18) & 63);
                }
                int q[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    q[c] = W.lat.get(cs[c]);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int sp = (del && cs[c] == sw) ? oldw : q[c];
                    ok &= (c >= nc) || (sp == ((cw[c] >> 24) & 0xFF));
                }
                for (int c = 4; c < nc; ++c) {
                    const int w = cd[c];
                    const int sc = site_from<IN>(W.d, lb, bx, by, bz, (w & 63) - 32, ((w >> 6) & 63) - 32,
                                                 ((w >> 12) & 63) - 32, (w >> 18) & 63);
                    const int sp = (del && sc == sw) ? oldw : W.lat.get(sc);
                    ok &= sp == ((w >> 24) & 0xFF);
                }
            }
            if (ok) {
                proc = arg;
                code = del ? MB_DEL_IF : MB_ADD;
            }
        } else {
            const int cx = wrap1(bx + dx, W.d.X);
            const int cy = W.d.dim >= 2 ? wrap1(by + dy, W.d.Y) : 0;
            const int cz = W.d.dim >= 3 ? wrap1(bz + dz, W.d.Z) : 0;
            s = W.d.lin(cx, cy, cz);
            proc = nli_eval<PPL, BITS, IN>(W, arg, cx, cy, cz);
            code = (kind == OP_DEL_NLI) ? MB_DEL_STRICT : MB_ADD;
        }
    }
    int slot = 0;
    if (nrounds <= 4) {
        // one result register per round: the rounds' atomics issue back to
        // back and are waited for once
        int sl[4] = {0, 0, 0, 0};
#pragma unroll
        for (int it = 0; it < 4; ++it)
            if (it < nrounds && proc > 0 && rnd == it)
                sl[it] = atomicAdd(&W.cnt[proc - 1], 1);
#pragma unroll
        for (int it = 0; it < 4; ++it)
            slot = (rnd == it) ? sl[it] : slot;
    } else {
        for (int it = 0; it < nrounds; ++it)
            if (proc > 0 && rnd == it)
                slot = atomicAdd(&W.cnt[proc - 1], 1);
    }
    if (proc > 0)
        W.mbox[slot * W.mstride + proc - 1] = ((s + 1) << 2) | code;
}

// Load op j of a segment and its first four condition words.
__device__ __forceinline__ void load_op(const int *prog, int ooff, int cbase, int j, bool valid, int2 &op,
                                        int (&cw)[4], const int *&cd)
{
    op = make_int2(0, 0);
    cd = prog;
    if (valid)
        op = *(const int2 *)(prog + ooff + OP_W * j);
    const int kind = op.x & 7;
    const int nc = (kind == OP_ADD_IF || kind == OP_DEL_IF) ? (op.x >> 17) & 0x3F : 0;
    cd = prog + cbase + (int)((unsigned)op.y >> 21);
    // unconditional reads (the program ends with 4 padding words)
#pragma unroll
    for (int c = 0; c < 4; ++c)
        cw[c] = cd[c];
#pragma unroll
    for (int c = 0; c < 4; ++c)
        cw[c] = (c < nc) ? cw[c] : 0;
}

// The static words a segment needs before it can run: its header and, per
// lane, the first op word and its first four condition words.  Read ahead
// of time (while the step waits for its site, or while the previous segment
// runs) to take them off the critical path.
struct SegPre {
    int nw, woff, nops, nrounds, ooff, cbase;
    bool as_add;
    int2 op;
    int cw[4];
    const int *cd;
};

__device__ __forceinline__ void seg_prefetch(const int *prog, int off_seg, int seg, int lane, SegPre &q)
{
    const int *sg = prog + off_seg + SEG_W * seg;
    q.nw = sg[0];
    q.woff = sg[1];
    q.nops = sg[2] & 0xFFFF;
    q.nrounds = (sg[2] >> 16) & 0xFF;
    q.as_add = (sg[2] >> 24) & 1;
    q.ooff = sg[3];
    q.cbase = sg[4];
    load_op(prog, q.ooff, q.cbase, lane, lane < q.nops, q.op, q.cw, q.cd);
}

// Post one segment at base cell (bx, by, bz) into the process queues: the
// lattice writes with the replace_species check (base.mpy:1187-1231), done
// by one lane per write, then the ops evaluated lane-parallel on the new
// lattice.  The op and condition words of the first 64 ops are read before
// the writes (they do not depend on the lattice).  Wave-uniform call.  On
// a species mismatch the writes before the failing one are done and
// nothing is posted.
template <int PPL, int BITS, bool IN>
__device__ __forceinline__ int post_segment_t(Wave<PPL, BITS> &W, const SegPre &q, int bx, int by, int bz)
{
    const int *prog = W.prog;
    const int nw = q.nw, woff = q.woff, nops = q.nops, nrounds = q.nrounds;
    const bool as_add = q.as_add;
    const int ooff = q.ooff, cbase = q.cbase;
    int2 op = q.op;
    int cw[4] = {q.cw[0], q.cw[1], q.cw[2], q.cw[3]};
    const int *cd = q.cd;
    const int lb = W.d.lin(bx, by, bz);

    // the site of a single write and its species before it (the lattice
    // the disable checks' conditions are tested on)
    int sw = -1, oldw = 0;
    if (nw > 0) {
        // the writes of a segment hit distinct sites (checked on the host)
        int s = 0, cur = 0, nsp = 0;
        bool bad = false;
        for (int w0i = 0; w0i < nw; w0i += 64) {
            const int w = w0i + W.lane;
            if (w < nw) {
                const int w0 = prog[woff + WR_W * w], w1 = prog[woff + WR_W * w + 1];
                s = site_from<IN>(W.d, lb, bx, by, bz, unpack_x(w0), unpack_y(w0), unpack_z(w0),
                                  (w0 >> 24) & 0xFF);
                cur = W.lat.get(s);
                nsp = w1 >> 16;
                bad = cur < 0 || !((w1 >> cur) & 1);
            }
            const unsigned long long bm = __ballot(w < nw && bad);
            const int nok = bm ? __ffsll((long long)bm) - 1 : 64;
            if (w < nw && W.lane < nok) {
                W.lat.xset(s, cur, nsp);
                if (W.track) {
                    // counts per species and site type (the null species of
                    // a multi-layer model is not counted)
                    const int t = prog[woff + WR_W * w] >> 24 & 0xFF;   // site type n
                    if (cur < W.S)
                        atomicAdd(&W.occ[cur * W.d.spuck + t - 1], -1);
                    if (nsp < W.S)
                        atomicAdd(&W.occ[nsp * W.d.spuck + t - 1], 1);
                }
            }
            if (bm) {
                wave_sync();
                return ST_MISMATCH;
            }
            if (nw == 1) {
                sw = readlane_i(s, 0);
                oldw = readlane_i(cur, 0);
            }
        }
        wave_sync();
    }

    eval_post<PPL, BITS, IN>(W, W.lane < nops, op, cw, cd, as_add, nrounds, lb, bx, by, bz, sw, oldw);
    for (int base = 64; base < nops; base += 64) {
        const int j = base + W.lane;
        load_op(prog, ooff, cbase, j, j < nops, op, cw, cd);
        eval_post<PPL, BITS, IN>(W, j < nops, op, cw, cd, as_add, nrounds, lb, bx, by, bz, sw, oldw);
    }
    wave_sync();
    return ST_OK;
}

// Two segments of one local_smart process posted in one pass (the common
// two-action process, e.g. a hop or a pair reaction).  Both writes are done
// first, then the ops of both segments are evaluated together, one op per
// lane; the first segment's ops must see the lattice before the second
// write, so their reads of the second write's site return its old species.
// Returns -1 (and changes nothing) when the pair does not qualify or a
// write would fail its species check: the caller then runs the segments
// one by one, which also reproduces the reference's behaviour on a
// mismatch.
template <int PPL, int BITS, bool IN, bool RA = false>
__device__ __forceinline__ int post_pair(Wave<PPL, BITS> &W, const SegPre &q1, const SegPre &q2, int bx1, int by1,
                                         int bz1, int bx2, int by2, int bz2, RePre &pre)
{
    const int *prog = W.prog;
    const int n1 = q1.nops, n2 = q2.nops;
    const int lb1 = W.d.lin(bx1, by1, bz1), lb2 = W.d.lin(bx2, by2, bz2);
    // the two writes, lanes 0 and 1
    const bool second = W.lane == 1;
    const int bx = second ? bx2 : bx1, by = second ? by2 : by1, bz = second ? bz2 : bz1;
    const int lbw = second ? lb2 : lb1;
    const int woff = second ? q2.woff : q1.woff;
    int s = 0, cur = 0, nsp = 0, t = 1;
    bool bad = false;
    if (W.lane < 2) {
        const int w0 = prog[woff], w1 = prog[woff + 1];
        t = (w0 >> 24) & 0xFF;
        s = site_from<IN>(W.d, lbw, bx, by, bz, unpack_x(w0), unpack_y(w0), unpack_z(w0), t);
        cur = W.lat.get(s);
        nsp = w1 >> 16;
        bad = cur < 0 || !((w1 >> cur) & 1);
    }
    const int sA = readlane_i(s, 0), sB = readlane_i(s, 1), oldB = readlane_i(cur, 1);
    const int oldA = readlane_i(cur, 0);
    if (__ballot(W.lane < 2 && bad) || sA == sB)
        return -1;
    if (W.lane < 2) {
        W.lat.xset(s, cur, nsp);
        if (W.track) {
            if (cur < W.S)
                atomicAdd(&W.occ[cur * W.d.spuck + t - 1], -1);
            if (nsp < W.S)
                atomicAdd(&W.occ[nsp * W.d.spuck + t - 1], 1);
        }
    }
    wave_sync();
    // the ops of both segments, one per lane
    const int j = W.lane;
    const bool valid = j < n1 + n2, first = j < n1;
    const int jj = first ? j : j - n1;
    const int ooff = first ? q1.ooff : q2.ooff, cbase = first ? q1.cbase : q2.cbase;
    const int ex = first ? bx1 : bx2, ey = first ? by1 : by2, ez = first ? bz1 : bz2;
    const int lb = first ? lb1 : lb2;
    int2 op;
    int cw[4];
    const int *cd;
    load_op(prog, ooff, cbase, jj, valid, op, cw, cd);
    int proc = 0, code = 0, so = 0, rnd = 0;
    if (valid) {
        const int kind = op.x & 7;
        rnd = ((op.x >> 3) & 0xFF) + (first ? 0 : q1.nrounds);
        const int n = (op.x >> 11) & 0x3F;
        const int arg = (int)((unsigned)op.x >> 23);
        const int dx = sext7(op.y & 0x7F), dy = sext7((op.y >> 7) & 0x7F), dz = sext7((op.y >> 14) & 0x7F);
        so = site_from<IN>(W.d, lb, ex, ey, ez, dx, dy, dz, n);
        const bool del = kind == OP_DEL_IF;
        const int nc = (op.x >> 17) & 0x3F;
        bool ok = true;
        // additions: the first segment's ops see site B before its write;
        // disable checks (lists matching the lattice): their process's
        // conditions on the lattice before their own segment's write, so
        // B is old for both and A old for the first segment's
        if (!del || (W.chk && nc > 0)) {
            int q[4], cs[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int w = (c < nc) ? cw[c] : (32 | 32 << 6 | 32 << 12 | 1 << 18);
                cs[c] = site_from<IN>(W.d, lb, ex, ey, ez, (w & 63) - 32, ((w >> 6) & 63) - 32,
                                      ((w >> 12) & 63) - 32, (w >> 18) & 63);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                q[c] = W.lat.get(cs[c]);
            const bool oldb = first || del, olda = first && del;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int sp = (oldb && cs[c] == sB) ? oldB : (olda && cs[c] == sA) ? oldA : q[c];
                ok &= (c >= nc) || (sp == ((cw[c] >> 24) & 0xFF));
            }
            for (int c = 4; c < nc; ++c) {
                const int w = cd[c];
                const int sc = site_from<IN>(W.d, lb, ex, ey, ez, (w & 63) - 32, ((w >> 6) & 63) - 32,
                                             ((w >> 12) & 63) - 32, (w >> 18) & 63);
                const int sp = (oldb && sc == sB) ? oldB : (olda && sc == sA) ? oldA : W.lat.get(sc);
                ok &= sp == ((w >> 24) & 0xFF);
            }
        }
        if (ok) {
            proc = arg;
            code = del ? MB_DEL_IF : MB_ADD;
        }
    }
    // RA: back[site] of a deletion, for the replay's first block (read now,
    // stored when the replay starts)
    const int bpre = (RA && PPL == 1 && proc > 0 && code != MB_ADD) ? W.back(proc - 1)[so + 1] : 0;
    const int nrt = q1.nrounds + q2.nrounds;
    int slot = 0;
    for (int it = 0; it < nrt; ++it)
        if (proc > 0 && rnd == it)
            slot = atomicAdd(&W.cnt[proc - 1], 1);
    if (proc > 0)
        W.mbox[slot * W.mstride + proc - 1] = ((so + 1) << 2) | code;
    if (RA && PPL == 1) {
        pre.qv = bpre;
        pre.qi = (proc > 0 && code != MB_ADD && slot < RBLK) ? slot * W.mstride + proc - 1 : -1;
    }
    wave_sync();
    return ST_OK;
}

// The static words of a lat_int group's one-pass posting, read while the
// step's site is on its way from HBM: whether the process qualifies (two
// segments at one offset, the first the deletions without writes, the
// second its alias with at most two writes), the item offset, each write
// lane's write words and each op lane's nli table header and coordinates.
struct LiPre {
    bool ok = false;
    int off, nw;
    int w0, w1;          // write lanes (lane < nw)
    int th, hdr;         // op lanes: the group's table and its header
    int cw[8];           // op lanes: the table's coordinate words
};

__device__ __forceinline__ void li_prefetch(const int *prog, int off_seg, int off_nlitab, int lane, int nit,
                                            const int *it, const SegPre &cur, LiPre &q)
{
    // every read stays inside the program whatever the process: a process
    // that does not qualify reads the program's first words instead
    const bool ok0 = nit == 2 && cur.nw == 0 && !cur.as_add && cur.nops <= 64;
    const int *ip = ok0 ? it : prog;
    const int i1 = ip[1], s2 = ip[IT_W], i2 = ip[IT_W + 1];
    const int *sg = prog + off_seg + SEG_W * (ok0 ? s2 : 0);
    const int nw2 = sg[0], woff2 = sg[1], h2 = sg[2], ooff2 = sg[3];
    q.ok = ok0 && i1 == i2 && nw2 <= 2 && ((h2 >> 24) & 1) && (h2 & 0xFFFF) == cur.nops && ooff2 == cur.ooff;
    q.off = i1;
    q.nw = q.ok ? nw2 : 0;
    const int wl = lane < q.nw ? lane : 0;
    q.w0 = prog[woff2 + WR_W * wl];
    q.w1 = prog[woff2 + WR_W * wl + 1];
    const int arg = lane < cur.nops ? (int)((unsigned)cur.op.x >> 23) : 0;
    q.th = prog[off_nlitab + arg];
    const int th = q.th < 0 ? 0 : q.th;
    q.hdr = prog[th];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        q.cw[k] = prog[th + 1 + k];
}

// The two segments of a lat_int group posted in one pass
// (run_proc_<group>, kmos/io/__init__.py:1497-1624): del_proc of every
// modified cell's process as nli_<group> finds it before the writes, the
// put/take writes, then add_proc as nli_<group> finds it after them.  Both
// segments hold the same ops, so each lane looks up its op's variant on
// both lattices from one read of its cluster: every read (the written
// sites and the clusters) sees the lattice before the writes, and the
// additions' lookups take the written sites' new species.  Every process's
// queue gets its deletions, then its additions, each in op order: the order
// of the two passes.  Returns -1 (and changes nothing) when a write
// would fail its species check: the caller then runs the segments one by
// one, which reproduces the reference's behaviour on a mismatch.  Every
// nli group of the model is tabulated.
template <int PPL, int BITS>
__device__ __forceinline__ int post_latint_pair(Wave<PPL, BITS> &W, const SegPre &q1, const LiPre &lp, int bx,
                                                int by, int bz, RePre &pre)
{
    const int *prog = W.prog;
    const int nops = q1.nops, R = q1.nrounds, nw = lp.nw;
    const int lb = W.d.lin(bx, by, bz);
    // the written sites (one lane each) and every op lane's cluster, read
    // together on the lattice before the writes
    const bool wr = W.lane < nw;
    const int t = (lp.w0 >> 24) & 0xFF;
    const int s = wr ? site_from<true>(W.d, lb, bx, by, bz, unpack_x(lp.w0), unpack_y(lp.w0), unpack_z(lp.w0), t)
                     : -1;
    const int nsp = lp.w1 >> 16;
    const int cur = W.lat.get(wr ? s : 0);
    const bool isop = W.lane < nops;
    const int2 op = q1.op;
    const int dx = sext7(op.y & 0x7F), dy = sext7((op.y >> 7) & 0x7F), dz = sext7((op.y >> 14) & 0x7F);
    const int cx = wrap1(bx + dx, W.d.X);
    const int cy = W.d.dim >= 2 ? wrap1(by + dy, W.d.Y) : 0;
    const int cz = W.d.dim >= 3 ? wrap1(bz + dz, W.d.Z) : 0;
    const int lc = W.d.lin(cx, cy, cz);
    const int K = lp.hdr & 0xFF, S = (lp.hdr >> 8) & 0xFF;
    int st[8], so[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int c = (isop && k < K) ? lp.cw[k] : (1 << 24);
        st[k] = site_from<true>(W.d, lc, cx, cy, cz, unpack_x(c), unpack_y(c), unpack_z(c), (c >> 24) & 0xFF);
        so[k] = W.lat.get(st[k]);
    }
    const bool bad = wr && (cur < 0 || !((lp.w1 >> cur) & 1));
    if (__ballot(bad))
        return -1;
    const int sA = readlane_i(s, 0), nA = readlane_i(nsp, 0);
    const int sB = readlane_i(s, 1), nB = readlane_i(nsp, 1);
    if (wr) {
        W.lat.xset(s, cur, nsp);
        if (W.track) {
            if (cur < W.S)
                atomicAdd(&W.occ[cur * W.d.spuck + t - 1], -1);
            if (nsp < W.S)
                atomicAdd(&W.occ[nsp * W.d.spuck + t - 1], 1);
        }
    }
    int pd = 0, pa = 0, rnd = -1;
    if (isop) {
        int i0 = 0, i1 = 0;
#pragma unroll
        for (int k = 7; k >= 0; --k) {
            const int sn = st[k] == sA ? nA : (st[k] == sB ? nB : so[k]);
            i0 = (k < K) ? i0 * S + so[k] : i0;
            i1 = (k < K) ? i1 * S + sn : i1;
        }
        const unsigned char *tab = (const unsigned char *)(prog + lp.th + 1 + K);
        pd = tab[i0];
        pa = tab[i1];
        rnd = (op.x >> 3) & 0xFF;
    }
    // back[site] of the deletion, for the replay's first block (read now,
    // stored when the replay starts)
    const int bpre = (PPL == 1 && pd > 0) ? W.back(pd - 1)[lc + 1] : 0;
    // the round of each pass's entry, -1 for none
    const int rd = pd > 0 ? rnd : -1, ra = pa > 0 ? rnd : -1;
    int sd = 0, sa = 0;
    if (R <= 8) {
        // queue slots from ranks (the queues are empty before the step's
        // only pass): a group's ops sit next to each other, at most 8, and
        // name processes no other group names, so the lanes within 7 of
        // this one hold every entry this lane's slots depend on; lanes of
        // other groups and lanes without an entry never match
        int lpd = pd, lpa = pa, rpd = pd, cda = (pa > 0 && pd == pa) ? 1 : 0;
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            lpd = __builtin_amdgcn_mov_dpp(lpd, 0x138, 0xF, 0xF, true);   // lane - k (0 below lane 0)
            lpa = __builtin_amdgcn_mov_dpp(lpa, 0x138, 0xF, 0xF, true);
            rpd = __builtin_amdgcn_mov_dpp(rpd, 0x130, 0xF, 0xF, true);   // lane + k (0 past lane 63)
            sd += (pd > 0 && lpd == pd) ? 1 : 0;
            sa += (pa > 0 && lpa == pa) ? 1 : 0;
            cda += (pa > 0 && lpd == pa) ? 1 : 0;
            cda += (pa > 0 && rpd == pa) ? 1 : 0;
        }
        // an addition follows every deletion of its process
        sa += cda;
        if (pd > 0)
            atomicAdd(&W.cnt[pd - 1], 1);
        if (pa > 0)
            atomicAdd(&W.cnt[pa - 1], 1);
    } else {
        for (int it = 0; it < R; ++it)
            if (rd == it)
                sd = atomicAdd(&W.cnt[pd - 1], 1);
        for (int it = 0; it < R; ++it)
            if (ra == it)
                sa = atomicAdd(&W.cnt[pa - 1], 1);
    }
    if (pd > 0)
        W.mbox[sd * W.mstride + pd - 1] = ((lc + 1) << 2) | MB_DEL_STRICT;
    if (pa > 0)
        W.mbox[sa * W.mstride + pa - 1] = ((lc + 1) << 2) | MB_ADD;
    if (PPL == 1) {
        pre.qv = bpre;
        pre.qi = (pd > 0 && sd < RBLK) ? sd * W.mstride + pd - 1 : -1;
    }
    wave_sync();
    return ST_OK;
}

// base cells away from the faces take the path without wrapping
template <int PPL, int BITS>
__device__ __forceinline__ int post_segment(Wave<PPL, BITS> &W, const SegPre &q, int bx, int by, int bz)
{
    if (W.d.interior(bx, by, bz))
        return post_segment_t<PPL, BITS, true>(W, q, bx, by, bz);
    return post_segment_t<PPL, BITS, false>(W, q, bx, by, bz);
}

// Execute one segment (a touchup) at base cell (bx, by, bz).
template <int PPL, int BITS>
__device__ int run_segment(Wave<PPL, BITS> &W, int seg, int bx, int by, int bz)
{
    SegPre q;
    seg_prefetch(W.prog, W.off_seg, seg, W.lane, q);
    const int e = post_segment<PPL, BITS>(W, q, bx, by, bz);
    NoMid none;
    RePre nopre;
    const int a = apply_queues<PPL, BITS>(W, none, nopre);
    return e ? e : a;
}

// The process's items and its first segment's static words, read ahead.
// The owner lanes' list slots nr, nr-1, ... (one process slot per lane),
// read before a post so that the first replay block need not read them.
template <int PPL, int BITS>
__device__ __forceinline__ void tail_ahead(const Wave<PPL, BITS> &W, RePre &q)
{
    const int nr0 = W.nr[0];
    const int *sl = W.sites(W.lane < W.P ? W.lane : 0);
#pragma unroll
    for (int j = 0; j < RBLK; ++j)
        q.tv[j] = sl[nr0 - j >= 1 ? nr0 - j : 0];
}

struct ProcPre {
    int nit;
    const int *it;
    SegPre seg;
};

__device__ __forceinline__ void proc_prefetch(const int *prog, int off_proc, int off_seg, int proc, int lane,
                                              ProcPre &q)
{
    const int *pp = prog + off_proc + 2 * (proc - 1);
    q.nit = pp[0];
    q.it = prog + pp[1];
    seg_prefetch(prog, off_seg, q.nit > 0 ? q.it[0] : 0, lane, q.seg);
}

// run_proc_nr(proc, site); procstat is counted by the caller.  With bit 0
// of `mode`, all segments of the process are queued and applied once, else
// each segment is applied before the next is posted.  LIP compiles the
// one-pass posting of a lat_int group's two segments, taken when bit 1 of
// `mode` is set (the model's nli groups are all tabulated).  The first segment's
// static words come read ahead in `pre`.
template <int PPL, int BITS, bool RB2 = false, class Mid = NoMid, bool LIP = false, bool RA = false>
__device__ int run_proc(Wave<PPL, BITS> &W, const ProcPre &pre, int site, int mode, Mid &mid,
                        const LiPre &li = LiPre{})
{
    const bool merged = mode & 1;
    const int s0 = site - 1;
    const int cell = divm(s0, W.ms);
    const int yz = divm(cell, W.mx);
    const int x = cell - yz * W.d.X;
    const int z = divm(yz, W.my);
    const int y = yz - z * W.d.Y;
    const int nit = pre.nit;
    const int *it = pre.it;
    SegPre cur = pre.seg;
    int e = ST_OK;
    // both segments posted in one pass (merged only): applied below
    bool posted = false;
    RePre rpre;       // read-ahead list words of the first replay block
    if (merged && W.local_smart && nit == 2 && cur.nw == 1 && !cur.as_add) {
        SegPre q2;
        seg_prefetch(W.prog, W.off_seg, it[IT_W], W.lane, q2);
        if (q2.nw == 1 && !q2.as_add && cur.nops + q2.nops <= 64) {
            const int o1 = it[1], o2 = it[IT_W + 1];
            const int bx1 = wrap1(x + unpack_x(o1), W.d.X);
            const int by1 = W.d.dim >= 2 ? wrap1(y + unpack_y(o1), W.d.Y) : 0;
            const int bz1 = W.d.dim >= 3 ? wrap1(z + unpack_z(o1), W.d.Z) : 0;
            const int bx2 = wrap1(x + unpack_x(o2), W.d.X);
            const int by2 = W.d.dim >= 2 ? wrap1(y + unpack_y(o2), W.d.Y) : 0;
            const int bz2 = W.d.dim >= 3 ? wrap1(z + unpack_z(o2), W.d.Z) : 0;
            // near the faces the segments run one by one
            if (W.d.interior(bx1, by1, bz1) && W.d.interior(bx2, by2, bz2)) {
                if (RA && PPL == 1)
                    tail_ahead(W, rpre);
                posted = post_pair<PPL, BITS, true, RA>(W, cur, q2, bx1, by1, bz1, bx2, by2, bz2, rpre) >= 0;
                rpre.on = RA && PPL == 1 && posted;
            }
        }
    }
    if (LIP && li.ok) {
        // a lat_int group: deletions, writes, additions over the same ops
        // (bit 1 is set only with bit 0; the static words in li)
        const int o = li.off;
        const int bx = wrap1(x + unpack_x(o), W.d.X);
        const int by = W.d.dim >= 2 ? wrap1(y + unpack_y(o), W.d.Y) : 0;
        const int bz = W.d.dim >= 3 ? wrap1(z + unpack_z(o), W.d.Z) : 0;
        if (W.d.interior(bx, by, bz)) {
            // the owner lanes' list tails, read while the pass posts
            if (PPL == 1)
                tail_ahead(W, rpre);
            posted = post_latint_pair<PPL, BITS>(W, cur, li, bx, by, bz, rpre) >= 0;
            rpre.on = PPL == 1 && posted;
        }
    }
    for (int i = 0; i < (posted ? 0 : nit); ++i, it += IT_W) {
        const int o = it[1];
        const int bx = wrap1(x + unpack_x(o), W.d.X);
        const int by = W.d.dim >= 2 ? wrap1(y + unpack_y(o), W.d.Y) : 0;
        const int bz = W.d.dim >= 3 ? wrap1(z + unpack_z(o), W.d.Z) : 0;
        if (i > 0)
            seg_prefetch(W.prog, W.off_seg, it[0], W.lane, cur);
        e = post_segment<PPL, BITS>(W, cur, bx, by, bz);
        if (e)
            break;
        if (!merged) {
            NoMid none;
            PH_MARK(W, 2);
            RePre nopre;
            e = apply_queues<PPL, BITS, RB2>(W, none, nopre);
            PH_MARK(W, 3);
            if (e)
                return e;
        }
    }
    PH_MARK(W, 2);
    const int a = merged ? apply_queues<PPL, BITS, RB2>(W, mid, rpre) : ST_OK;
    PH_MARK(W, 3);
    return e ? e : a;
}

// lane l receives v from lane l-1 (DPP wave_shr:1); lane 0 receives 0
__device__ __forceinline__ double wave_shr1z(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}


// accum_rates (base.mpy:603-623) as a strictly left-to-right FP64 sum; the
// partial sum of process p ends on lane p%64, slot p/64.  Each pass shifts
// the partial sums one lane up (lane 0 receives the sum of the slots
// before) and adds every lane's own term: after pass i lane i holds the
// exact sequential prefix ((c0 + c1) + c2) + ..., the same rounding as the
// reference's loop, and a lane that holds its final prefix keeps it in
// later passes (its left neighbour is final too).  Prefixes below the first
// process whose term changed (`first`, per slot) are kept from the last
// step, so only the passes from there on run.  Returns the total.
template <int PPL>
__device__ __forceinline__ double accumulate(const double (&c)[PPL], double (&acc)[PPL], int P, int lane,
                                             const int (&first)[PPL])
{
    double carry = 0.0;
    bool dirty = false;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int lim = min(64, P - 64 * k);
        if (lim <= 0) {
            acc[k] = 0.0;
            continue;
        }
        const int f = dirty ? 0 : first[k];
        if (f < lim) {
            // lane 0's term carries the sum of the slots before (shifted-in 0)
            const double ce = (lane == 0) ? carry + c[k] : c[k];
            double a = (lane >= f) ? ce : acc[k];
            // four passes per loop trip: one branch per four links of the
            // chain instead of one per link
            int it = max(f, 1);
            for (; it + 4 <= lim; it += 4) {
                a = wave_shr1z(a) + ce;
                a = wave_shr1z(a) + ce;
                a = wave_shr1z(a) + ce;
                a = wave_shr1z(a) + ce;
            }
            for (; it < lim; ++it)
                a = wave_shr1z(a) + ce;
            acc[k] = a;
            dirty = true;
        }
        carry = readlane_d(acc[k], lim - 1);
    }
    return carry;
}

// interval_search_real (base.mpy:1234-1338) on the lane-distributed
// accum_rates: first index with value < accum (else the last), then the
// leftmost index holding the same accum value.  Returns a 0-based index.
template <int PPL>
__device__ __forceinline__ int interval_search(const double (&acc)[PPL], double value, int P, int lane)
{
    int istar = P - 1;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const unsigned long long mk = __ballot((lane + 64 * k < P) && (value < acc[k]));
        if (mk) {
            istar = 64 * k + __ffsll((long long)mk) - 1;
            break;
        }
    }
    const double target = readlane_d(slot_get<PPL>(acc, istar >> 6), istar & 63);
    int pick = istar;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const unsigned long long mk = __ballot((lane + 64 * k < P) && (acc[k] == target));
        if (mk) {
            pick = 64 * k + __ffsll((long long)mk) - 1;
            break;
        }
    }
    return pick;
}

// The next step's accumulate and search, run as the replay's hook in the
// plain step loop (one process slot per lane) while the replay's HBM reads
// are on their way.  From the count each lane's process will have if every
// deletion of the block finds its site registered, it updates this lane's
// prefix in place and gives the next step's total and pick; the caller
// keeps them only when the replay left exactly those counts, and otherwise
// recomputes every prefix.
struct NextSel {
    static constexpr bool active = true;
    bool armed = false;
    bool on;             // the next step's u_proc is at hand
    double u1;           // the next step's u_proc
    double &acc;         // this lane's prefix (PPL == 1)
    int nrs = 0;         // predicted count of this lane's process
    double total = 0.0;
    int pick = 0;
    bool made = false;   // prefixes were changed
    __device__ NextSel(bool on_, double u1_, double &acc_) : on(on_), u1(u1_), acc(acc_) {}
    template <class Wv>
    __device__ __forceinline__ void operator()(const Wv &W, int nr_after)
    {
        nrs = nr_after;
        if (!on)
            return;
        const double c[1] = {(double)nr_after * W.rate[0]};
        const unsigned long long ch = __ballot(nr_after != W.nr[0]);
        const int fs[1] = {ch ? __ffsll((long long)ch) - 1 : 64};
        double a[1] = {acc};
        total = accumulate<1>(c, a, W.P, W.lane, fs);
        acc = a[0];
        if (total > 0.0)
            pick = interval_search<1>(a, u1 * total, W.P, W.lane);
        made = true;
    }
};

// Copy the op program to LDS (whole workgroup), then set up this wave.
// Returns the replica index, or -1 for a wave without one.
template <int PPL, int BITS, bool LDS_LAT>
__device__ __forceinline__ int wave_setup(Wave<PPL, BITS> &W, const KParams &kp, unsigned char *smem)
{
    int *sprog = (int *)smem;
    for (int i = threadIdx.x; i < kp.prog_words; i += blockDim.x)
        sprog[i] = kp.prog[i];
    __syncthreads();
    // wave-uniform values: keep them in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = blockIdx.x * kp.wpg + wave;
    if (slot >= kp.nrep)
        return -1;
    const int r = __builtin_amdgcn_readfirstlane(kp.rep_ids[slot]);
    unsigned char *wsm = smem + kp.prog_lds + (size_t)wave * kp.wave_lds;
    W.prog = sprog;
    W.lane = threadIdx.x & 63;
    W.P = kp.P;
    W.S = kp.S;
    W.mstride = 64 * PPL;
    W.off_seg = kp.off_seg;
    W.off_proc = kp.off_proc;
    W.off_touch = kp.off_touch;
    W.off_nli = kp.off_nli;
    W.off_nlitab = kp.off_nlitab;
    const int4 dd0 = kp.dims[r];
    int4 dd;
    dd.x = __builtin_amdgcn_readfirstlane(dd0.x);
    dd.y = __builtin_amdgcn_readfirstlane(dd0.y);
    dd.z = __builtin_amdgcn_readfirstlane(dd0.z);
    dd.w = __builtin_amdgcn_readfirstlane(dd0.w);
    W.d.X = dd.x; W.d.Y = dd.y; W.d.Z = dd.z; W.d.V = dd.w;
    W.d.spuck = kp.spuck; W.d.dim = kp.dim; W.d.moff = kp.moff;
    W.d.strides();
    W.mx = make_magic(dd.x);
    W.my = make_magic(dd.y);
    W.ms = make_magic(kp.spuck);
    W.V1 = (long long)dd.w + 1;
    W.avail = kp.avail_pool + kp.avail_off[r];
    unsigned char *glat = kp.lat_pool + kp.lat_off[r];
    if (LDS_LAT) {
        const long long lb = lattice_bytes(dd.w, BITS);
        const uint4 *src = (const uint4 *)glat;
        uint4 *dst = (uint4 *)wsm;
        for (long long i = W.lane; i < lb / 16; i += 64)
            dst[i] = src[i];
        W.lat.p = wsm;
        W.mbox = (int *)(wsm + kp.lat_lds);
    } else {
        W.lat.p = glat;
        W.mbox = (int *)wsm;
    }
    W.mrows = kp.mrows;
    W.cnt = W.mbox + kp.mrows * W.mstride;
    W.win = W.cnt + W.mstride;
    W.occ = W.win + 64 * WIN;
    W.bvq = W.occ + ((kp.S * kp.spuck + 3) & ~3);
    W.track = false;
    W.chk = false;
    W.local_smart = kp.backend == 0;
#if KMCG_PHASES
    for (int i = 0; i < 5; ++i)
        W.tph[i] = 0;
    W.nblk = W.napply = 0;
    W.tmark = __builtin_amdgcn_s_memtime();
#endif
    for (int i = W.lane; i < W.mstride; i += 64)
        W.cnt[i] = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = W.lane + 64 * k;
        if (p < W.P) {
            const long long o = (long long)r * W.P + p;
            W.nr[k] = kp.nr[o];
            W.rate[k] = kp.rates[o];
            W.integ[k] = kp.integ[o];
            W.pstat[k] = kp.pstat[o];
        } else {
            W.nr[k] = 0; W.rate[k] = 0.0; W.integ[k] = 0.0; W.pstat[k] = 0;
        }
    }
    wave_sync();
    return r;
}

template <int PPL, int BITS, bool LDS_LAT>
__device__ __forceinline__ void wave_store(Wave<PPL, BITS> &W, const KParams &kp, int r, bool lattice_changed)
{
    wave_sync();
    if (LDS_LAT && lattice_changed) {
        const long long lb = lattice_bytes(W.d.V, BITS);
        uint4 *dst = (uint4 *)(kp.lat_pool + kp.lat_off[r]);
        const uint4 *src = (const uint4 *)W.lat.p;
        for (long long i = W.lane; i < lb / 16; i += 64)
            dst[i] = src[i];
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = W.lane + 64 * k;
        if (p < W.P) {
            const long long o = (long long)r * W.P + p;
            kp.nr[o] = W.nr[k];
            kp.integ[o] = W.integ[k];
            kp.pstat[o] = W.pstat[k];
        }
    }
}

// One sample record (the layout of kmc_sample_kernel) from the wave's
// registers and its species counts.
template <int PPL, int BITS>
__device__ __forceinline__ void write_record(const Wave<PPL, BITS> &W, const KParams &kp, int slot, long long idx,
                                             double t, double dt, long long ks, int st)
{
    if (idx < 0 || idx >= kp.srec_cap)
        return;
    wave_sync();
    double *rec = kp.srec + ((long long)slot * kp.srec_cap + idx) * kp.srec_width;
    const int no = kp.S * W.d.spuck, P = W.P;
    const double cells = (double)W.d.X * W.d.Y * W.d.Z;
    for (int i = W.lane; i < no; i += 64)
        rec[4 + i] = (double)W.occ[i] / cells;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = W.lane + 64 * k;
        if (p < P) {
            rec[4 + no + p] = W.integ[k];
            rec[4 + no + P + p] = (double)W.pstat[k];
        }
    }
    if (W.lane == 0) {
        rec[0] = t;
        rec[1] = dt;
        rec[2] = (double)ks;
        rec[3] = (double)st;
    }
}

// do_kmc_steps for one replica per wave (proclist_generic_subroutines.mpy:1-44)
// Displacement tracking of one step (do_kmc_steps_displacement,
// proclist_generic_subroutines_acf.mpy:143-219): get_diff_sites_displacement
// for the chosen (process, site) from the hop table (the site pair and the
// displacement kmos/io/acf.py:443-671 generates), then assign_particle_id,
// update_displacement, update_id_arr and, when a trajectory is recorded,
// the step's event (base_acf.mpy:825-891).  A process that is no hop moves
// no particle.  Lane 0 does the work.
template <int PPL, int BITS>
__device__ __forceinline__ void msd_track(const Wave<PPL, BITS> &W, const KParams &kp, int r, int proc0, int site)
{
    if (W.lane != 0)
        return;
    const int *h = kp.msd_hop + MSD_HOP_W * proc0;
    int p = 0, fin = 0;
    if (h[0] && kp.acf_pa) {
        // do_kmc_steps_acf: the sites move, the displacement does not
        // (get_diff_sites_acf, assign_particle_id, update_id_arr)
        const int s0 = site - 1;
        const int cell = divm(s0, W.ms);
        const int yz = divm(cell, W.mx);
        const int x = cell - yz * W.d.X;
        const int z = divm(yz, W.my);
        const int y = yz - z * W.d.Y;
        const int init = site0(W.d, x, y, z, h[1], h[2], h[3], h[4]) + 1;
        fin = site0(W.d, x, y, z, h[5], h[6], h[7], h[8]) + 1;
        const long long o = kp.msd_off[r];
        int *id = kp.msd_id + o;
        int *sa = kp.msd_site + o;
        p = id[init];
        id[init] = 0;
        id[fin] = p;
        sa[p] = fin;
        sa[0] = 0;
    } else if (h[0]) {
        const int s0 = site - 1;
        const int cell = divm(s0, W.ms);
        const int yz = divm(cell, W.mx);
        const int x = cell - yz * W.d.X;
        const int z = divm(yz, W.my);
        const int y = yz - z * W.d.Y;
        const int init = site0(W.d, x, y, z, h[1], h[2], h[3], h[4]) + 1;
        fin = site0(W.d, x, y, z, h[5], h[6], h[7], h[8]) + 1;
        const long long o = kp.msd_off[r];
        int *id = kp.msd_id + o;
        int *sa = kp.msd_site + o;
        p = id[init];
        double *dp = kp.msd_disp + 3 * (o + p);
        const double *d = kp.msd_d + 3 * proc0;
        dp[0] = dp[0] + d[0];
        dp[1] = dp[1] + d[1];
        dp[2] = dp[2] + d[2];
        id[init] = 0;
        id[fin] = p;
        sa[p] = fin;
        sa[0] = 0;
    }
    if (kp.acf_pa) {
        kp.acf_i[8 * r + 5] = p;            // the particle update_buffer_acf takes
        kp.acf_i[8 * r + 6] = fin;
        return;                              // its trajectory event follows run_proc_nr
    }
    if (kp.msd_evcap > 0) {
        const long long k = kp.msd_evn[r];
        if (k < kp.msd_evcap) {
            int *ev = kp.msd_ev + 2 * ((long long)r * kp.msd_evcap + k);
            ev[0] = p;
            ev[1] = fin;
            kp.msd_evn[r] = k + 1;
        }
    }
}

// update_config_bin (base_acf.mpy:1098-1172): the step's time, weighted by
// sum(buffer_acf) over nr_of_ions_before_kmc_step_acf, spread over the time
// bins since t_o.  Lane 0.  A step longer than the extended bins reaches
// stops at the last bin (the reference writes past its array there).
__device__ __forceinline__ void acf_bins(const KParams &kp, int r, double t, double dt)
{
    double *d = kp.acf_d + 8 * r;
    int *iv = kp.acf_i + 8 * r;
    double *cb = kp.acf_bin + kp.acf_boff[r] - 1;     // 1-based like config_bin
    int *cc = kp.acf_cnt + kp.acf_boff[r] - 1;
    const double sum = d[0];
    const double entry = (sum == 0.0) ? 0.0 : sum / (double)iv[3];
    const double t_o = d[1], t_bin = d[3];
    const int nb = iv[1], ext = iv[2];
    int bin = iv[0];
    if (t - t_o < (double)(bin + 1) * t_bin) {
        cb[bin + 1] = cb[bin + 1] + entry * dt;
        cc[bin + 1] += 1;
    } else {
        cb[bin + 1] = cb[bin + 1] + entry * ((double)(bin + 1) * t_bin - (t - t_o - dt));
        cc[bin + 1] += 1;
        bin += 1;
        if (bin < nb) {
            while ((double)(bin + 1) * t_bin < t - t_o && bin + 1 < ext) {
                cb[bin + 1] = cb[bin + 1] + entry * t_bin;
                cc[bin + 1] += 1;
                bin += 1;
            }
            cb[bin + 1] = cb[bin + 1] + entry * (t - t_o - (double)bin * t_bin);
            cc[bin + 1] += 1;
        }
    }
    iv[0] = bin;
}

// update_property_and_buffer_acf (base_acf.mpy:1075-1096), whole wave: each
// particle's property now and buffer_acf = product_property(p, p); the sum
// of buffer_acf is exact in any order (the products are integers).
template <int PPL, int BITS>
__device__ __forceinline__ void acf_refresh(const Wave<PPL, BITS> &W, const KParams &kp, int r)
{
    const long long o = kp.msd_off[r];
    const int *sa = kp.msd_site + o;
    const int *pa = kp.acf_pa + o;
    int *po = kp.acf_po + o;
    double *buf = kp.acf_buf + o;
    const int nt1 = kp.acf_nt1;
    double part = 0.0;
    for (int i = W.lane; i <= W.d.V; i += 64) {
        if (i >= 1) {
            const int q = pa[sa[i]];
            po[i] = q;
            buf[i] = kp.acf_pp[q * nt1 + q];
        }
        part += buf[i];
    }
    for (int sh = 32; sh >= 1; sh >>= 1)
        part += __shfl_xor(part, sh);
    if (W.lane == 0)
        kp.acf_d[8 * r] = part;
}

// The ACF steps of do_kmc_steps_acf before run_proc_nr (update_config_bin,
// then the tracking of msd_track, then update_after_wrap_acf,
// base_acf.mpy:1174-1201).  Wave-uniform call.
template <int PPL, int BITS>
__device__ __forceinline__ void acf_before(const Wave<PPL, BITS> &W, const KParams &kp, int r, double t,
                                           double dt, int proc0, int site)
{
    if (W.lane == 0)
        acf_bins(kp, r, t, dt);
    msd_track<PPL, BITS>(W, kp, r, proc0, site);
    int wrap = 0;
    if (W.lane == 0) {
        int *iv = kp.acf_i + 8 * r;
        if (iv[0] > iv[1] - 1) {
            double *d = kp.acf_d + 8 * r;
            d[1] = t;
            iv[0] = 0;
            d[2] = d[2] + 1.0;
            iv[3] = iv[4];
            wrap = 1;
        }
    }
    if (__builtin_amdgcn_readfirstlane(wrap))
        acf_refresh<PPL, BITS>(W, kp, r);
}

// update_buffer_acf (base_acf.mpy:1053-1073) after run_proc_nr, then the
// step's trajectory event.  Lane 0.
__device__ __forceinline__ void acf_after(const KParams &kp, int r)
{
    const int p = kp.acf_i[8 * r + 5], fin = kp.acf_i[8 * r + 6];
    const long long o = kp.msd_off[r];
    double *buf = kp.acf_buf + o;
    const double nv = kp.acf_pp[kp.acf_po[o + p] * kp.acf_nt1 + kp.acf_pa[o + kp.msd_site[o + p]]];
    kp.acf_d[8 * r] = kp.acf_d[8 * r] + (nv - buf[p]);
    buf[p] = nv;
    if (kp.msd_evcap > 0) {
        const long long k = kp.msd_evn[r];
        if (k < kp.msd_evcap) {
            int *ev = kp.msd_ev + 2 * ((long long)r * kp.msd_evcap + k);
            ev[0] = p;
            ev[1] = fin;
            kp.msd_evn[r] = k + 1;
        }
    }
}

// OV (one process slot per lane, no tracking): the next step's accumulate
// and search run during the replay's HBM reads.  It shortens a trajectory's
// chain of dependent steps and adds instructions, so it pays where few
// replica-waves share a SIMD (latency-bound) and not where many do.
template <int PPL, int BITS, bool LDS_LAT, bool MSD, bool OV, bool LIP>
__device__ __forceinline__ void steps_loop(const KParams &kp, unsigned char *smem)
{
    Wave<PPL, BITS> W;
    const int r = wave_setup<PPL, BITS, LDS_LAT>(W, kp, smem);
    if (r < 0)
        return;
    const int slot = blockIdx.x * kp.wpg + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = W.lane, P = W.P;
    // a replica that has stopped takes no step (and still gives its record)
    const int st0 = kp.status[r];
    W.chk = kp.lists_ok && kp.lists_ok[r] && W.local_smart;
    long long n = kp.nsteps;
    if (kp.budget)
        n = min(n, kp.budget[slot]);
    if (st0 != ST_OK)
        n = 0;
    if (kp.srec) {
        // species counts per site type, then kept up to date by the writes
        W.track = true;
        const int no = kp.S * W.d.spuck;
        for (int i = lane; i < no; i += 64)
            W.occ[i] = 0;
        wave_sync();
        for (int s = lane; s < W.d.V; s += 64) {
            const int q = W.lat.get(s);
            if (q >= 0 && q < W.S)
                atomicAdd(&W.occ[q * W.d.spuck + (s - divm(s, W.ms) * W.d.spuck)], 1);
        }
        wave_sync();
    }
    double t = kp.ktime[r], dt_last = kp.kdt[r];
    long long ks = kp.kstep[r];
    unsigned long long ctr = kp.rngctr[r];
    const unsigned seed = kp.seed[r], stream = kp.stream[r];
    int st = st0;
    long long done = 0;
    // records (while `track`): one every sample_every steps, counted down
    // instead of a 64-bit remainder per step
    int rec_left = kp.sample_every;
    long long rec_idx = 0;
    // accum_rates stay in registers across steps; `first` is the lowest
    // lane per slot whose term changed (64 = none), everything at the start
    double acc[PPL];
    int first[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        acc[k] = 0.0;
        first[k] = 0;
    }
    // uniforms of 64 steps at a time: lane i draws step ubase + i (its
    // -log(u_time), u_proc, u_site), each step reads its lane's values
    double lmlog, lu1, lu2;
    unsigned long long ubase = ctr;
    int uj = 0;
    {
        double uu[3];
        step_uniforms(seed, stream, ubase + lane, uu);
        lmlog = -log(uu[0]);
        lu1 = uu[1];
        lu2 = uu[2];
    }
    // the next step's total and pick, when the replay computed them (one
    // process slot per lane, no tracking)
    bool have = false;
    double s_total = 0.0;
    int s_pick = 0;
#if KMCG_PHASES
    for (int i = 0; i < 5; ++i)
        W.tph[i] = 0;
    W.nblk = W.napply = 0;
    W.tmark = __builtin_amdgcn_s_memtime();
#endif
    while (done < n) {
        PH_MARK(W, 4);
        if (uj == 64) {
            ubase += 64;
            uj = 0;
            double uu[3];
            step_uniforms(seed, stream, ubase + lane, uu);
            lmlog = -log(uu[0]);
            lu1 = uu[1];
            lu2 = uu[2];
        }
        const double mlog = readlane_d(lmlog, uj);
        double u[3];
        u[1] = readlane_d(lu1, uj);
        u[2] = readlane_d(lu2, uj);
        double c[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            c[k] = (double)W.nr[k] * W.rate[k];
        const double total = have ? s_total : accumulate<PPL>(c, acc, P, lane, first);
        if (!(total > 0.0)) {
            st = ST_ABSORBING;
            break;
        }
        ctr += 1;
        // update_clocks (base.mpy:1123-1161)
        const double dt = mlog / total;
        t = t + dt;
        dt_last = dt;
        ks += 1;
        // update_integ_rate (base.mpy:626-645): (nr*rate)*dt, pre-step nr
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            W.integ[k] = W.integ[k] + c[k] * dt;
        // determine_procsite (base.mpy:1075-1120)
        const int pick = have ? s_pick : interval_search<PPL>(acc, u[1] * total, P, lane);
        PH_MARK(W, 0);
        const int pk = pick >> 6, pl = pick & 63;
        const int nrp = readlane_i(slot_get<PPL>(W.nr, pk), pl);
        int kk = (int)(1.0 + u[2] * (double)nrp);
        kk = kk > nrp ? nrp : kk;
        if (kk < 1) {                       // only a corrupted database gets here
            st = ST_CORRUPT;
            break;
        }
        int sv = 0;
        if (lane == pl) {
            sv = W.sites(pick)[kk];
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (k == pk) W.pstat[k] += 1;
        }
        // the process's program is read while its site is on the way
        ProcPre pre;
        proc_prefetch(W.prog, W.off_proc, W.off_seg, pick + 1, lane, pre);
        // (the LIP form runs only models with bit 1 of kp.merged)
        LiPre li;
        if constexpr (LIP)
            li_prefetch(W.prog, W.off_seg, W.off_nlitab, lane, pre.nit, pre.it, pre.seg, li);
        uj += 1;
        const int site = readlane_i(sv, pl);
        PH_MARK(W, 1);
        done += 1;
        if (MSD) {
            if (kp.acf_pa)
                acf_before<PPL, BITS>(W, kp, r, t, dt, pick, site);
            else
                msd_track<PPL, BITS>(W, kp, r, pick, site);
        }
        int nr_before[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            nr_before[k] = W.nr[k];
        bool redo = false;
        if constexpr (OV && PPL == 1 && !MSD) {
            NextSel nx(uj < 64, readlane_d(lu1, uj & 63), acc[0]);
            // (no read-ahead of the replay's list words here: their HBM
            // round trip is what hides the next step's select; measured
            // slower with read-ahead and the select run before its wait,
            // local_smart and lat_int alike)
            st = run_proc<PPL, BITS, true>(W, pre, site, kp.merged, nx);
            have = nx.made && !st && !__ballot(W.nr[0] != nx.nrs);
            redo = nx.made && !have;
            s_total = nx.total;
            s_pick = nx.pick;
        } else {
            NoMid none;
            st = run_proc<PPL, BITS, !MSD, NoMid, LIP, PPL == 1 && !MSD>(W, pre, site, kp.merged, none, li);
        }
        if (MSD && kp.acf_pa && lane == 0)
            acf_after(kp, r);
        if (st)
            break;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const unsigned long long ch = __ballot(W.nr[k] != nr_before[k]);
            // prefixes the replay's hook changed on a wrong guess: all anew
            first[k] = redo ? 0 : ch ? __ffsll((long long)ch) - 1 : 64;
        }
        if (W.track && --rec_left == 0) {
            write_record<PPL, BITS>(W, kp, slot, rec_idx, t, dt_last, ks, ST_OK);
            rec_left = kp.sample_every;
            rec_idx += 1;
        }
    }
    if (W.track)
        write_record<PPL, BITS>(W, kp, slot, done == 0 ? 0 : (done - 1) / kp.sample_every, t, dt_last, ks, st);
    wave_store<PPL, BITS, LDS_LAT>(W, kp, r, true);
#if KMCG_PHASES
    PH_MARK(W, 4);
    if (lane == 0 && kp.phase) {
        for (int i = 0; i < 5; ++i)
            kp.phase[PHASE_W * r + i] += W.tph[i];
        kp.phase[PHASE_W * r + 5] += (unsigned long long)done;
        kp.phase[PHASE_W * r + 6] += W.nblk;
        kp.phase[PHASE_W * r + 7] += W.napply;
    }
#endif
    if (lane == 0) {
        kp.ktime[r] = t;
        kp.kdt[r] = dt_last;
        kp.kstep[r] = ks;
        kp.rngctr[r] = ctr;
        kp.status[r] = st;
        kp.executed[slot] = done;
        if (kp.budget)
            kp.budget[slot] -= done;
    }
}

template <int PPL, int BITS, bool LDS_LAT, bool MSD, bool OV = false>
__global__ __launch_bounds__(64 * MAX_WPG) __attribute__((amdgpu_waves_per_eu(4))) void kmc_steps_kernel(KParams kp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    steps_loop<PPL, BITS, LDS_LAT, MSD, OV, false>(kp, smem);
}

// The LIP form (lat_int models whose nli groups are all tabulated): a
// group's two segments posted in one pass, their static words read while
// the step's site is on its way.
template <int PPL, int BITS, bool LDS_LAT>
__global__ __launch_bounds__(64 * MAX_WPG) __attribute__((amdgpu_waves_per_eu(4))) void kmc_steps_kernel_li(KParams kp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    steps_loop<PPL, BITS, LDS_LAT, false, false, true>(kp, smem);
}

// touchups, run_proc_nr, get_next_kmc_step (one replica per wave)
template <int PPL, int BITS, bool LDS_LAT>
__global__ __launch_bounds__(64 * MAX_WPG) void kmc_misc_kernel(KParams kp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    Wave<PPL, BITS> W;
    const int r = wave_setup<PPL, BITS, LDS_LAT>(W, kp, smem);
    if (r < 0)
        return;
    const int lane = W.lane, P = W.P;
    const int *prog = W.prog;
    int st = kp.status[r];
    bool changed = false;
    if (kp.mode == MODE_INIT || kp.mode == MODE_ADJUST || kp.mode == MODE_TOUCHUP) {
        const bool lat_int = kp.backend == 1;
        const int nsite = lat_int ? 1 : W.d.spuck;
        if (kp.mode == MODE_INIT) {
            for (int z = 0; z < W.d.Z && !st; ++z)
                for (int y = 0; y < W.d.Y && !st; ++y)
                    for (int x = 0; x < W.d.X && !st; ++x)
                        for (int n = 1; n <= nsite && !st; ++n)
                            st = run_segment<PPL, BITS>(W, prog[W.off_touch + n - 1], x, y, z);
        } else if (kp.mode == MODE_ADJUST) {
            for (int x = 0; x < W.d.X && !st; ++x)
                for (int y = 0; y < W.d.Y && !st; ++y)
                    for (int z = 0; z < W.d.Z && !st; ++z)
                        for (int n = 1; n <= nsite && !st; ++n)
                            st = run_segment<PPL, BITS>(W, prog[W.off_touch + n - 1], x, y, z);
        } else {
            const int seg = prog[W.off_touch + (lat_int ? 0 : kp.arg3 - 1)];
            st = run_segment<PPL, BITS>(W, seg, kp.arg0, kp.arg1, kp.arg2);
        }
    } else if (kp.mode == MODE_RUN_PROC) {
        const int proc = kp.arg0, site = kp.arg1;
        const int pk = (proc - 1) >> 6, pl = (proc - 1) & 63;
        if (lane == pl) {
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (k == pk) W.pstat[k] += 1;
        }
        ProcPre pre;
        proc_prefetch(W.prog, W.off_proc, W.off_seg, proc, lane, pre);
        NoMid none;
        st = run_proc<PPL, BITS>(W, pre, site, kp.merged, none);
        changed = true;
    } else if (kp.mode == MODE_NEXT_STEP) {
        double c[PPL], acc[PPL];
        int first[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            c[k] = (double)W.nr[k] * W.rate[k];
            acc[k] = 0.0;
            first[k] = 0;
        }
        const double total = accumulate<PPL>(c, acc, P, lane, first);
        if (!(total > 0.0)) {
            st = ST_ABSORBING;
        } else {
            double u[3];
            const unsigned long long ctr = kp.rngctr[r];
            step_uniforms(kp.seed[r], kp.stream[r], ctr, u);
            const int pick = interval_search<PPL>(acc, u[1] * total, P, lane);
            const int pk = pick >> 6, pl = pick & 63;
            const int nrp = readlane_i(slot_get<PPL>(W.nr, pk), pl);
            int kk = (int)(1.0 + u[0] * (double)nrp);
            kk = kk > nrp ? nrp : kk;
            int sv = 0;
            if (lane == pl && kk >= 1)
                sv = W.sites(pick)[kk];
            const int site = readlane_i(sv, pl);
            if (lane == 0) {
                kp.out[0] = pick + 1;
                kp.out[1] = site;
                kp.rngctr[r] = ctr + 1;
            }
        }
    }
    wave_store<PPL, BITS, LDS_LAT>(W, kp, r, changed);
    if (lane == 0)
        kp.status[r] = st;
}

// default species of every site (initialize_state); site defaults are the
// first spuck words of the program
template <int BITS>
__global__ void kmc_fill_default_kernel(KParams kp)
{
    const int r = kp.rep_ids[blockIdx.x];
    const int4 dd = kp.dims[r];
    unsigned char *lat = kp.lat_pool + kp.lat_off[r];
    const int *sdef = kp.prog;
    const long long lb = lattice_bytes(dd.w, BITS);
    if (BITS == 8) {
        for (long long s = threadIdx.x; s < lb; s += blockDim.x)
            lat[s] = (unsigned char)(signed char)(s < dd.w ? sdef[s % kp.spuck] : 0);
    } else {
        const int per = 8 / BITS;
        for (long long b = threadIdx.x; b < lb; b += blockDim.x) {
            unsigned v = 0;
            for (int i = 0; i < per; ++i) {
                const long long s = b * per + i;
                const unsigned code = s < dd.w ? (unsigned)(sdef[s % kp.spuck] + 1) : 0u;
                v |= code << (i * BITS);
            }
            lat[b] = (unsigned char)v;
        }
    }
}

// nli_<group>(cell) on the program in HBM (initialisation kernel).
template <int BITS>
__device__ int nli_eval_g(const int *prog, int off_nli, int group, const Lattice<BITS> &lat,
                          const Dims &d, int cx, int cy, int cz)
{
    int pos = prog[off_nli + group];
    for (;;) {
        const int *nd = prog + pos;
        const int nc = nd[1];
        if (nc == -1)
            return nd[2];
        const int c = nd[0];
        const int sp = lat.get(site0(d, cx, cy, cz, unpack_x(c), unpack_y(c), unpack_z(c), (c >> 24) & 0xFF));
        int next = -1;
        for (int k = 0; k < nc; ++k) {
            const int *e = nd + 2 + 3 * k;
            if (sp >= 0 && ((e[0] >> sp) & 1)) {
                if (e[1] == 1)
                    return e[2];
                next = e[2];
                break;
            }
        }
        if (next < 0)
            return 0;
        pos = next;
    }
}

// initialize_state on an empty database (proclist_generic_subroutines.mpy:
// 236-304).  The touchups, run in k, j, i order, append every (process,
// site) pair at most once and in increasing site order, so the list of each
// process is the ordered set of sites where it is enabled.  One workgroup
// builds one process's list of one replica by a block-wide compaction.
// init entry per process (3 words at kp.arg0): site type of its anchor (bit
// 16: a site type initialize_state does not touch up, bit 17: multi-layer
// model), the word offset of its touchup addition op (local_smart, -1:
// never enabled) or its group (lat_int), the op's condition base.
template <int BITS>
__global__ __launch_bounds__(256) void kmc_init_lists_kernel(KParams kp)
{
    __shared__ int wave_tot[4];
    const int P = kp.P;
    const int slot = blockIdx.x / P, p = blockIdx.x % P;
    const int r = kp.rep_ids[slot];
    const int4 dd = kp.dims[r];
    Dims d;
    d.X = dd.x; d.Y = dd.y; d.Z = dd.z; d.V = dd.w; d.spuck = kp.spuck; d.dim = kp.dim; d.moff = kp.moff;
    d.strides();
    Lattice<BITS> lat;
    lat.p = kp.lat_pool + kp.lat_off[r];
    const int *prog = kp.prog;
    const int *ie = prog + kp.arg0 + 3 * p;
    const int n_exec = ie[0] & 0xFFFF, arg = ie[1], cbase = ie[2];
    const bool lat_int = kp.backend == 1;
    if (!lat_int && (arg < 0 || ((ie[0] >> 16) & 1)))
        return;
    const long long V1 = (long long)dd.w + 1;
    int *sp = kp.avail_pool + kp.avail_off[r] + (long long)p * V1;
    int *bp = kp.avail_pool + kp.avail_off[r] + (long long)(P + p) * V1;
    const int ncell = dd.x * dd.y * dd.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int w0 = 0, w1 = 0;
    if (!lat_int) {
        w0 = prog[arg];
        w1 = prog[arg + 1];
    }
    const int nc = (w0 >> 17) & 0x3F;
    const int *cd = prog + cbase + (int)((unsigned)w1 >> 21);
    int base = 0;
    for (int t = 0; t < ncell; t += 256) {
        const int c = t + threadIdx.x;
        bool en = false;
        if (c < ncell) {
            const int x = c % dd.x, y = (c / dd.x) % dd.y, z = c / (dd.x * dd.y);
            if (lat_int) {
                en = nli_eval_g<BITS>(prog, kp.off_nli, arg, lat, d, x, y, z) == p + 1;
            } else {
                en = true;
                for (int k = 0; k < nc && en; ++k) {
                    const int cw = cd[k];
                    en = lat.get(site0(d, x, y, z, (cw & 63) - 32, ((cw >> 6) & 63) - 32,
                                       ((cw >> 12) & 63) - 32, (cw >> 18) & 63)) == ((cw >> 24) & 0xFF);
                }
            }
        }
        const unsigned long long mask = __ballot(en);
        if (lane == 0)
            wave_tot[wave] = __popcll(mask);
        __syncthreads();
        int before = base;
        for (int w = 0; w < wave; ++w)
            before += wave_tot[w];
        const int total = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        if (en) {
            const int pos = before + __popcll(mask & ((1ull << lane) - 1ull)) + 1;
            const int site = kp.spuck * c + n_exec;
            sp[pos] = site;
            bp[site] = pos;
        }
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        kp.nr[(long long)r * P + p] = base;
}

// The debug check of one replica's lists (kmcg_check_lists), one block per
// (replica, process): the invariant add_proc/del_proc keep (base.mpy:
// 211-301: sites[1..nr] distinct sites of the process's executing site type
// with back[site] = position, every other slot and back entry 0), and the
// registered set against a scan of every cell with the process's
// conditions (local_smart: the touchup's if-tree conditions; lat_int: the
// nli case tree), the set kmc_init_lists_kernel builds.  Adds the number of
// violations to bad[slot].
template <int BITS>
__global__ __launch_bounds__(256) void kmc_check_lists_kernel(KParams kp, int *bad)
{
    const int P = kp.P;
    const int slot = blockIdx.x / P, p = blockIdx.x % P;
    const int r = kp.rep_ids[slot];
    const int4 dd = kp.dims[r];
    Dims d;
    d.X = dd.x; d.Y = dd.y; d.Z = dd.z; d.V = dd.w; d.spuck = kp.spuck; d.dim = kp.dim; d.moff = kp.moff;
    d.strides();
    Lattice<BITS> lat;
    lat.p = kp.lat_pool + kp.lat_off[r];
    const int *prog = kp.prog;
    const int *ie = prog + kp.arg0 + 3 * p;
    const int n_exec = ie[0] & 0xFFFF, arg = ie[1], cbase = ie[2];
    const bool lat_int = kp.backend == 1;
    // multi-layer: processes anchored on layers the initialisation does
    // not touch up can hold while unregistered, as in the reference
    const bool subset = (ie[0] >> 17) & 1;
    const int V = dd.w;
    const long long V1 = (long long)V + 1;
    const int *sp = kp.avail_pool + kp.avail_off[r] + (long long)p * V1;
    const int *bp = kp.avail_pool + kp.avail_off[r] + (long long)(P + p) * V1;
    const int nr = kp.nr[(long long)r * P + p];
    const bool never = !lat_int && arg < 0;      // no condition set: never enabled
    int nbad = (nr < 0 || nr > V || (never && nr != 0)) ? 1 : 0;
    for (int k = threadIdx.x + 1; k <= V; k += blockDim.x) {
        const int site = sp[k];
        if (k <= nr)
            nbad += (site < 1 || site > V || bp[site] != k || (site - 1) % d.spuck + 1 != n_exec) ? 1 : 0;
        else
            nbad += site != 0 ? 1 : 0;
        const int back = bp[k];
        if (back != 0)
            nbad += (back < 0 || back > nr || sp[back] != k) ? 1 : 0;
    }
    if (!never) {
        int w0 = 0, w1 = 0;
        if (!lat_int) {
            w0 = prog[arg];
            w1 = prog[arg + 1];
        }
        const int nc = (w0 >> 17) & 0x3F;
        const int *cd = prog + cbase + (int)((unsigned)w1 >> 21);
        const int ncell = dd.x * dd.y * dd.z;
        for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
            const int x = c % dd.x, y = (c / dd.x) % dd.y, z = c / (dd.x * dd.y);
            bool en;
            if (lat_int) {
                en = nli_eval_g<BITS>(prog, kp.off_nli, arg, lat, d, x, y, z) == p + 1;
            } else {
                en = true;
                for (int k = 0; k < nc && en; ++k) {
                    const int cw = cd[k];
                    en = lat.get(site0(d, x, y, z, (cw & 63) - 32, ((cw >> 6) & 63) - 32,
                                       ((cw >> 12) & 63) - 32, (cw >> 18) & 63)) == ((cw >> 24) & 0xFF);
                }
            }
            const bool reg = bp[kp.spuck * c + n_exec] != 0;
            nbad += (subset ? reg && !en : en != reg) ? 1 : 0;
        }
    }
    if (nbad)
        atomicAdd(&bad[slot], nbad);
}

// get_occupation (proclist_generic_subroutines.mpy:113-158): counts per
// (species, site type); out[slot][S][spuck] as fractions of the cell count.
template <int BITS>
__global__ void kmc_occupation_kernel(KParams kp, double *out)
{
    __shared__ int counts[8 * 32];
    const int r = kp.rep_ids[blockIdx.x];
    const int4 dd = kp.dims[r];
    const int S = kp.S, spuck = kp.spuck;
    for (int i = threadIdx.x; i < S * spuck; i += blockDim.x)
        counts[i] = 0;
    __syncthreads();
    Lattice<BITS> lat;
    lat.p = kp.lat_pool + kp.lat_off[r];
    for (int s = threadIdx.x; s < dd.w; s += blockDim.x) {
        const int sp = lat.get(s);
        if (sp >= 0 && sp < S)
            atomicAdd(&counts[sp * spuck + s % spuck], 1);
    }
    __syncthreads();
    const double cells = (double)dd.x * dd.y * dd.z;
    for (int i = threadIdx.x; i < S * spuck; i += blockDim.x)
        out[(long long)blockIdx.x * S * spuck + i] = (double)counts[i] / cells;
}

// One sampling record per listed replica (kmos/run/__init__.py:680-850
// reads the same quantities through get_atoms): kmc_time, kmc_time_step,
// kmc_step, status, occupation[S][spuck] (per cell), integ_rates[P],
// procstat[P], all as doubles.  One workgroup per replica.
template <int BITS>
__global__ __launch_bounds__(256) void kmc_sample_kernel(KParams kp, double *out, int width)
{
    __shared__ int counts[8 * 32];
    const int r = kp.rep_ids[blockIdx.x];
    const int4 dd = kp.dims[r];
    const int S = kp.S, spuck = kp.spuck, P = kp.P;
    for (int i = threadIdx.x; i < S * spuck; i += blockDim.x)
        counts[i] = 0;
    __syncthreads();
    Lattice<BITS> lat;
    lat.p = kp.lat_pool + kp.lat_off[r];
    for (int s = threadIdx.x; s < dd.w; s += blockDim.x) {
        const int sp = lat.get(s);
        if (sp >= 0 && sp < S)
            atomicAdd(&counts[sp * spuck + s % spuck], 1);
    }
    __syncthreads();
    double *rec = out + (long long)blockIdx.x * width;
    const double cells = (double)dd.x * dd.y * dd.z;
    for (int i = threadIdx.x; i < S * spuck; i += blockDim.x)
        rec[4 + i] = (double)counts[i] / cells;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        rec[4 + S * spuck + p] = kp.integ[(long long)r * P + p];
        rec[4 + S * spuck + P + p] = (double)kp.pstat[(long long)r * P + p];
    }
    if (threadIdx.x == 0) {
        rec[0] = kp.ktime[r];
        rec[1] = kp.kdt[r];
        rec[2] = (double)kp.kstep[r];
        rec[3] = (double)kp.status[r];
    }
}

} // namespace

// ===========================================================================
// host code
// ===========================================================================
namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(KMCG_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int NO_WRITE = -1000;

struct OpH {
    int kind, round, arg, dx, dy, dz, n;
    std::vector<int> conds;   // 5 ints per condition
};
struct SegH {
    std::vector<int> writes;  // 6 ints per write
    std::vector<OpH> ops;
    int nrounds = 0;
    int alias_of = -1;        // lat_int additions: the ops of this deletion segment
};

} // namespace

struct kmcg_model {
    std::vector<int> blob;
    std::vector<int> prog;   // packed op program (layout in kmcg_device.h)
    int P = 0, S = 0, spuck = 0, dim = 0, backend = 0, nullsp = -1, maxoff = 0, maxR = 1;
    int moff = 0;        // largest offset component any segment reaches
    int rstep = 1;       // mailbox rows of one executed process (all its segments)
    int mrows = 1;       // mailbox rows per wave
    int merged = 0;      // bit 0: one apply per step; bit 1: lat_int groups in one pass
    int off_seg = 0, off_proc = 0, off_touch = 0, off_nli = 0, off_init = 0, off_nlitab = 0;
    // a species id a lattice site may hold: 0..S-1, or the null species S of
    // a multi-layer model (sites of layers that are not populated)
    bool species_ok(int v) const { return (v >= 0 && v < S) || (nullsp >= 0 && v == nullsp); }
    // lattice bits per site: 2 when every value, null included, fits
    // species+1 in 0..3 (single layer: null -1 -> 0), else 8
    int lattice_bits() const { return (nullsp >= 0 ? S + 1 : S) + 1 <= 4 ? 2 : 8; }
};

struct kmcg_batch {
    const kmcg_model *m = nullptr;
    int device = 0;
    int n = 0;
    int bits = 8;
    int lds_mode = -1;
    int max_wpg = MAX_WPG;
    int ov_mode = -1;      // overlapped step kernel: -1 automatic, 0 off, 1 on
    int n_cu = 256;        // compute units of the batch's device
    int max_lds = 65536;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<int> X, Y, Z, V;
    std::vector<long long> lat_off, avail_off;
    long long lat_total = 0, avail_total = 0;
    // device buffers
    int *prog = nullptr;
    int4 *dims = nullptr;
    long long *lat_off_d = nullptr, *avail_off_d = nullptr;
    unsigned *seed = nullptr, *strm = nullptr;
    unsigned char *lat_pool = nullptr;
    int *avail_pool = nullptr;
    int *nr = nullptr;
    double *rates = nullptr, *integ = nullptr;
    long long *pstat = nullptr;
    double *ktime = nullptr, *kdt = nullptr;
    long long *kstep = nullptr;
    unsigned long long *rngctr = nullptr;
    int *status = nullptr;
    long long *executed = nullptr;
    long long *budget = nullptr;
    int *rep_ids = nullptr;
    int *out = nullptr;
    unsigned long long *phase = nullptr;   // phase counters (KMCG_PHASES builds)
    double *occ = nullptr;
    double *samples = nullptr;
    // pinned staging of a run's replica list and budgets; one run in flight
    int *h_ids = nullptr;
    long long *h_budget = nullptr;
    int64_t *user_budget = nullptr;    // where the budgets go back at the sync
    double *srec = nullptr;             // in-loop records on the device
    double *h_srec = nullptr;           // and their pinned staging
    size_t srec_len = 0;                // doubles in each
    double *user_srec = nullptr;
    size_t n_srec = 0;
    double *h_samples = nullptr;        // pinned staging of sample records
    double *user_samples = nullptr;     // where the records go at the sync
    size_t n_samples = 0;
    int n_last = 0;
    float last_ms = 0.f;
    bool timed_run_pending = false;   // ev0/ev1 bracket a launch not yet read
    // kmcg_set_graph: a run's stream work (replica list and budget copies,
    // the step kernel, budget and record copies back) captured in a HIP
    // graph, keyed by everything the captured nodes hold, and replayed
    // while a run repeats it; the copies read the pinned staging buffers,
    // so a replay moves their contents at the time it runs
    struct Captured {
        std::vector<unsigned char> key;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    bool use_graph = false;
    std::vector<Captured> graphs;     // oldest first, at most MAX_GRAPHS
    // 1 per replica whose lists match its lattice (set by kmcg_init and
    // kmcg_set_config, cleared by kmcg_replace_species and kmcg_set_state)
    int *lists_ok = nullptr;
    // displacement tracking (kmcg_msd_*): allocated by the first kmcg_msd_init
    int *msd_hop = nullptr;
    double *msd_d = nullptr;
    int *msd_id = nullptr, *msd_site = nullptr;
    double *msd_disp = nullptr;
    long long *msd_off_d = nullptr;
    std::vector<long long> msd_off;
    std::vector<int> msd_nions;         // tracked particles per replica (0: not initialised)
    int *msd_ev = nullptr;
    long long *msd_evn = nullptr;
    long long msd_evcap = 0;
    // autocorrelation bins (kmcg_acf_*): allocated by kmcg_acf_alloc
    int *acf_pa = nullptr, *acf_po = nullptr;
    double *acf_buf = nullptr, *acf_pp = nullptr;
    int acf_nt1 = 0;
    double *acf_bin = nullptr;
    int *acf_cnt = nullptr;
    long long *acf_boff = nullptr;
    double *acf_d = nullptr;
    int *acf_i = nullptr;
    int acf_ext = 0;
};

namespace {

// Rounds: an op runs after every earlier op whose possible processes
// intersect its own (a static process, or the members of an nli group).
void assign_rounds(SegH &sg, const std::vector<std::vector<int>> &group_procs, int P)
{
    const int W = (P + 63) / 64;
    std::vector<std::vector<unsigned long long>> sets;
    for (auto &op : sg.ops) {
        std::vector<unsigned long long> b(W, 0ull);
        if (op.kind == OP_DEL_IF || op.kind == OP_ADD_IF) {
            b[(op.arg - 1) >> 6] |= 1ull << ((op.arg - 1) & 63);
        } else {
            for (int p : group_procs[op.arg])
                b[(p - 1) >> 6] |= 1ull << ((p - 1) & 63);
        }
        sets.push_back(b);
    }
    sg.nrounds = 0;
    for (size_t j = 0; j < sg.ops.size(); ++j) {
        int rk = 0;
        for (size_t i = 0; i < j; ++i) {
            bool inter = false;
            for (int w = 0; w < W; ++w)
                inter |= (sets[i][w] & sets[j][w]) != 0;
            if (inter)
                rk = std::max(rk, sg.ops[i].round + 1);
        }
        sg.ops[j].round = rk;
        sg.nrounds = std::max(sg.nrounds, rk + 1);
    }
}

int build_prog(kmcg_model *m)
{
    const std::vector<int> &b = m->blob;
    auto W = [&](long long i) -> int {
        if (i < 0 || i >= (long long)b.size())
            throw std::runtime_error("blob offset out of range");
        return b[i];
    };
    try {
        m->backend = W(2);
        m->P = W(3);
        m->S = W(4);
        m->nullsp = W(5);
        m->spuck = W(6);
        m->dim = W(7);
        m->maxoff = W(15);
        if (m->backend != 0 && m->backend != 1)
            return fail(KMCG_EINVAL, "unknown backend in model blob");
        if (m->P < 1 || m->P > 256)
            return fail(KMCG_EUNSUPPORTED, "engine supports 1..256 processes");
        if (m->S < 1 || m->S > 8)
            return fail(KMCG_EUNSUPPORTED, "engine supports 1..8 species");
        if (m->spuck < 1 || m->dim < 1 || m->dim > 3)
            return fail(KMCG_EINVAL, "bad spuck/dimension");
        if (m->nullsp != -1 && m->nullsp != m->S)
            return fail(KMCG_EINVAL, "null species must be -1 (one layer) or nr_of_species");

        std::vector<SegH> segs;
        std::vector<std::vector<int>> proc_items(m->P);   // 4 ints per item
        std::vector<int> touchup;
        std::vector<std::vector<int>> group_procs;
        std::vector<std::vector<int>> nli_nodes;            // per group, relative offsets
        if (m->backend == 0) {
            const int nrt = W(10), rtab = W(11);
            for (int r = 0; r < nrt; ++r) {
                const int base = rtab + 7 * r;
                SegH sg;
                const int site_type = W(base), old = W(base + 1), nw = W(base + 2);
                if (old != NO_WRITE) {
                    if (!m->species_ok(old) || !m->species_ok(nw))
                        return fail(KMCG_EINVAL, "bad species in routine");
                    sg.writes = {0, 0, 0, site_type, 1 << old, nw};
                }
                const int ndis = W(base + 3), doff = W(base + 4);
                for (int i = 0; i < ndis; ++i) {
                    const int o = doff + 5 * i;
                    sg.ops.push_back({OP_DEL_IF, 0, W(o), W(o + 1), W(o + 2), W(o + 3), W(o + 4), {}});
                }
                const int nen = W(base + 5), eoff = W(base + 6);
                for (int i = 0; i < nen; ++i) {
                    const int o = eoff + 7 * i;
                    OpH op{OP_ADD_IF, 0, W(o), W(o + 1), W(o + 2), W(o + 3), W(o + 4), {}};
                    const int nc = W(o + 5), co = W(o + 6);
                    for (int c = 0; c < nc; ++c)
                        for (int q = 0; q < 5; ++q)
                            op.conds.push_back(W(co + 5 * c + q));
                    sg.ops.push_back(op);
                }
                segs.push_back(sg);
            }
            const int ptab = W(9);
            for (int p = 0; p < m->P; ++p) {
                const int na = W(ptab + 3 * p + 1), ao = W(ptab + 3 * p + 2);
                for (int a = 0; a < na; ++a) {
                    const int o = ao + 5 * a;
                    proc_items[p].insert(proc_items[p].end(), {W(o), W(o + 1), W(o + 2), W(o + 3)});
                }
            }
            for (int n = 0; n < m->spuck; ++n)
                touchup.push_back(W(W(12) + n));
            // every process's conditions, relative to its executing site's
            // cell: the enable of the touchup routine of that site type
            std::vector<std::vector<int>> pconds(m->P + 1);
            std::vector<bool> have(m->P + 1, false);
            for (int n = 0; n < m->spuck; ++n) {
                const int t = touchup[n];
                if (t < 0 || t >= (int)segs.size())
                    return fail(KMCG_EINVAL, "bad touchup routine");
                for (auto &op : segs[t].ops)
                    if (op.kind == OP_ADD_IF && op.arg >= 1 && op.arg <= m->P && !have[op.arg]) {
                        pconds[op.arg] = op.conds;
                        have[op.arg] = true;
                    }
            }
            // a disable check of a put/take routine carries the conditions of
            // its process at the checked site: while the lists match the
            // lattice, the process is registered there exactly when they
            // hold on the lattice before the routine's write, so the check
            // needs no read of avail_sites when they fail
            std::vector<bool> is_touchup(segs.size(), false);
            for (int t : touchup)
                is_touchup[t] = true;
            for (size_t si = 0; si < segs.size(); ++si) {
                if (is_touchup[si] || segs[si].writes.size() != 6)
                    continue;
                for (auto &op : segs[si].ops) {
                    if (op.kind != OP_DEL_IF || op.arg < 1 || op.arg > m->P || !have[op.arg])
                        continue;
                    const std::vector<int> &pc = pconds[op.arg];
                    for (size_t c = 0; c < pc.size(); c += 5)
                        op.conds.insert(op.conds.end(),
                                        {op.dx + pc[c], op.dy + pc[c + 1], op.dz + pc[c + 2], pc[c + 3], pc[c + 4]});
                }
            }
        } else {
            const int G = W(13), gtab = W(14);
            for (int g = 0; g < G; ++g) {
                const int base = gtab + 8 * g;
                std::vector<int> procs;
                for (int i = 0; i < W(base + 6); ++i)
                    procs.push_back(W(W(base + 7) + i));
                group_procs.push_back(procs);
                std::vector<int> nodes(b.begin() + W(base + 4), b.begin() + W(base + 4) + W(base + 5));
                nli_nodes.push_back(nodes);
            }
            for (int g = 0; g < G; ++g) {
                const int base = gtab + 8 * g;
                SegH del, add;
                const int nmod = W(base), moff = W(base + 1);
                for (int i = 0; i < nmod; ++i) {
                    const int o = moff + 4 * i;
                    del.ops.push_back({OP_DEL_NLI, 0, W(o), W(o + 1), W(o + 2), W(o + 3), 1, {}});
                    add.ops.push_back({OP_ADD_NLI, 0, W(o), W(o + 1), W(o + 2), W(o + 3), 1, {}});
                }
                // the ops of one nli group next to each other, in their
                // order (ops of different groups name disjoint processes, so
                // every process's queue keeps the reference order): the
                // one-pass posting ranks a group's entries within its block
                auto by_group = [](const OpH &a, const OpH &b) { return a.arg < b.arg; };
                std::stable_sort(del.ops.begin(), del.ops.end(), by_group);
                std::stable_sort(add.ops.begin(), add.ops.end(), by_group);
                const int nupd = W(base + 2), uoff = W(base + 3);
                for (int i = 0; i < nupd; ++i)
                    for (int q = 0; q < 6; ++q)
                        add.writes.push_back(W(uoff + 6 * i + q));
                add.alias_of = (int)segs.size();
                segs.push_back(del);
                segs.push_back(add);
            }
            for (int p = 0; p < m->P; ++p) {
                const int g = W(W(9) + p);
                if (g < 0 || g >= G)
                    return fail(KMCG_EINVAL, "bad group index");
                proc_items[p] = {2 * g, 0, 0, 0, 2 * g + 1, 0, 0, 0};
            }
            SegH tu;
            for (int p = 1; p <= m->P; ++p)
                tu.ops.push_back({OP_DEL_IF, 0, p, 0, 0, 0, 1, {}});
            for (int g = 0; g < G; ++g)
                tu.ops.push_back({OP_ADD_NLI, 0, g, 0, 0, 0, 1, {}});
            touchup.push_back((int)segs.size());
            segs.push_back(tu);
        }
        m->maxR = 1;
        for (auto &sg : segs) {
            for (auto &op : sg.ops) {
                if ((op.kind == OP_DEL_IF || op.kind == OP_ADD_IF) && (op.arg < 1 || op.arg > m->P))
                    return fail(KMCG_EINVAL, "op names a process out of range");
                if (op.n < 1 || op.n > m->spuck)
                    return fail(KMCG_EINVAL, "op names a site type out of range");
            }
            assign_rounds(sg, group_procs, m->P);
            m->maxR = std::max(m->maxR, sg.nrounds);
        }

        // the writes of one segment hit distinct sites (they run in parallel)
        for (auto &sg : segs) {
            const int nw = (int)sg.writes.size() / 6;
            for (int a = 0; a < nw; ++a)
                for (int c = a + 1; c < nw; ++c)
                    if (std::equal(sg.writes.begin() + 6 * a, sg.writes.begin() + 6 * a + 4,
                                   sg.writes.begin() + 6 * c))
                        return fail(KMCG_EUNSUPPORTED, "a routine writes one site twice");
        }
        // rows of one executed process; merge its segments if they fit 8 KiB
        m->rstep = 1;
        for (int p = 0; p < m->P; ++p) {
            int rows = 0;
            for (size_t i = 0; i < proc_items[p].size(); i += 4) {
                const int sgi = proc_items[p][i];
                if (sgi >= 0 && sgi < (int)segs.size())
                    rows += segs[sgi].nrounds;
            }
            m->rstep = std::max(m->rstep, rows);
        }
        {
            const int ppl = m->P <= 64 ? 1 : m->P <= 128 ? 2 : 4;
            m->merged = (size_t)m->rstep * 64 * ppl * sizeof(int) <= 8192 ? 1 : 0;
            m->mrows = m->merged ? std::max(m->rstep, m->maxR) : m->maxR;
        }
        // deletions of a process precede its additions in every segment
        for (auto &sg : segs) {
            bool seen_add = false;
            for (auto &op : sg.ops) {
                const bool add = op.kind == OP_ADD_IF || op.kind == OP_ADD_NLI;
                if (!add && seen_add)
                    return fail(KMCG_EINVAL, "segment deletes after it adds");
                seen_add |= add;
            }
        }
        auto fits = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
        // pack (layout in kmcg_device.h)
        std::vector<int> &pg = m->prog;
        pg.clear();
        for (int n = 0; n < m->spuck; ++n)
            pg.push_back(W(W(8) + n));
        auto align4 = [&]() { while (pg.size() % 4) pg.push_back(0); };
        align4();
        m->off_seg = (int)pg.size();
        pg.resize(pg.size() + SEG_W * segs.size(), 0);
        for (size_t si = 0; si < segs.size(); ++si) {
            SegH &sg = segs[si];
            if (sg.ops.size() > 0xFFFF || sg.nrounds > 255)
                return fail(KMCG_EUNSUPPORTED, "segment too large");
            const int nw = (int)sg.writes.size() / 6;
            const int sb = m->off_seg + SEG_W * (int)si;
            pg[sb + 0] = nw;
            pg[sb + 1] = (int)pg.size();
            for (int w = 0; w < nw; ++w) {
                const int *wr = &sg.writes[6 * w];
                if (!fits(wr[0], -127, 127) || !fits(wr[1], -127, 127) || !fits(wr[2], -127, 127))
                    return fail(KMCG_EUNSUPPORTED, "offset too large");
                pg.push_back(pack3(wr[0], wr[1], wr[2]) | (wr[3] << 24));
                pg.push_back((wr[4] & 0xFFFF) | (wr[5] << 16));
            }
            if (sg.alias_of >= 0) {
                // additions of a lat_int group: same op list as its deletions
                const SegH &src = segs[sg.alias_of];
                bool same = src.ops.size() == sg.ops.size() && src.nrounds == sg.nrounds;
                for (size_t j = 0; same && j < sg.ops.size(); ++j)
                    same = src.ops[j].kind == OP_DEL_NLI && sg.ops[j].kind == OP_ADD_NLI &&
                           src.ops[j].arg == sg.ops[j].arg && src.ops[j].round == sg.ops[j].round &&
                           src.ops[j].dx == sg.ops[j].dx && src.ops[j].dy == sg.ops[j].dy &&
                           src.ops[j].dz == sg.ops[j].dz;
                if (!same || sg.alias_of >= (int)si)
                    return fail(KMCG_EINVAL, "lat_int addition segment does not mirror its deletions");
                const int ab = m->off_seg + SEG_W * sg.alias_of;
                pg[sb + 2] = (int)sg.ops.size() | (sg.nrounds << 16) | (1 << 24);
                pg[sb + 3] = pg[ab + 3];
                pg[sb + 4] = pg[ab + 4];
                continue;
            }
            pg[sb + 2] = (int)sg.ops.size() | (sg.nrounds << 16);
            if (pg.size() % 2)
                pg.push_back(0);
            const int ob = (int)pg.size();
            pg[sb + 3] = ob;
            pg.resize(pg.size() + OP_W * sg.ops.size(), 0);
            const int cb = (int)pg.size();
            pg[sb + 4] = cb;
            for (size_t j = 0; j < sg.ops.size(); ++j) {
                const OpH &op = sg.ops[j];
                const int o = ob + OP_W * (int)j;
                const int nc = (int)op.conds.size() / 5;
                const int crel = (int)pg.size() - cb;
                if (!fits(op.dx, -63, 63) || !fits(op.dy, -63, 63) || !fits(op.dz, -63, 63) ||
                    nc > 63 || op.round > 255 || op.n > 63 || op.arg < 0 || op.arg > 511 || crel > 2047)
                    return fail(KMCG_EUNSUPPORTED, "op out of packing range");
                pg[o + 0] = op.kind | (op.round << 3) | (op.n << 11) | (nc << 17) | (op.arg << 23);
                pg[o + 1] = (op.dx & 127) | ((op.dy & 127) << 7) | ((op.dz & 127) << 14) | (crel << 21);
                for (int c = 0; c < nc; ++c) {
                    const int *cd = &op.conds[5 * c];
                    if (!fits(cd[0], -32, 31) || !fits(cd[1], -32, 31) || !fits(cd[2], -32, 31) ||
                        cd[3] > 63 || cd[4] < 0 || cd[4] > 255)
                        return fail(KMCG_EUNSUPPORTED, "condition out of packing range");
                    pg.push_back((cd[0] + 32) | ((cd[1] + 32) << 6) | ((cd[2] + 32) << 12) |
                                 (cd[3] << 18) | (cd[4] << 24));
                }
            }
        }
        m->off_proc = (int)pg.size();
        pg.resize(pg.size() + 2 * m->P, 0);
        for (int p = 0; p < m->P; ++p) {
            pg[m->off_proc + 2 * p] = (int)proc_items[p].size() / 4;
            pg[m->off_proc + 2 * p + 1] = (int)pg.size();
            for (size_t i = 0; i < proc_items[p].size(); i += 4) {
                const int *it = &proc_items[p][i];
                if (it[0] < 0 || it[0] >= (int)segs.size())
                    return fail(KMCG_EINVAL, "process program names a bad segment");
                if (!fits(it[1], -127, 127) || !fits(it[2], -127, 127) || !fits(it[3], -127, 127))
                    return fail(KMCG_EUNSUPPORTED, "offset too large");
                pg.push_back(it[0]);
                pg.push_back(pack3(it[1], it[2], it[3]));
            }
        }
        for (int t : touchup)
            if (t < 0 || t >= (int)segs.size())
                return fail(KMCG_EINVAL, "bad touchup segment");
        m->off_touch = (int)pg.size();
        pg.insert(pg.end(), touchup.begin(), touchup.end());
        m->off_nli = (int)pg.size();
        pg.resize(pg.size() + nli_nodes.size(), 0);
        for (size_t g = 0; g < nli_nodes.size(); ++g) {
            // blob nodes: [dx, dy, dz, n, n_cases, (mask, kind, target)*] or
            // [.., -1, value]; children at blob-relative offsets
            const std::vector<int> &src = nli_nodes[g];
            std::vector<int> old_pos, new_pos;
            size_t pos = 0, npos = pg.size();
            while (pos < src.size()) {
                if (pos + 5 > src.size())
                    return fail(KMCG_EINVAL, "truncated nli node");
                const int nc = src[pos + 4];
                old_pos.push_back((int)pos);
                new_pos.push_back((int)npos);
                if (nc == -1) {
                    if (pos + 6 > src.size())
                        return fail(KMCG_EINVAL, "truncated nli leaf");
                    pos += 6;
                    npos += 3;
                } else {
                    if (nc < 0 || pos + 5 + 3 * (size_t)nc > src.size())
                        return fail(KMCG_EINVAL, "bad nli node");
                    pos += 5 + 3 * nc;
                    npos += 2 + 3 * nc;
                }
            }
            pg[m->off_nli + g] = (int)pg.size();
            for (size_t q = 0; q < old_pos.size(); ++q) {
                const int *nd = &src[old_pos[q]];
                const int nc = nd[4];
                if (nc == -1) {
                    if (nd[5] < 0 || nd[5] > m->P)
                        return fail(KMCG_EINVAL, "bad nli leaf");
                    pg.insert(pg.end(), {0, -1, nd[5]});
                    continue;
                }
                if (nd[3] < 1 || nd[3] > m->spuck || !fits(nd[0], -127, 127) ||
                    !fits(nd[1], -127, 127) || !fits(nd[2], -127, 127))
                    return fail(KMCG_EINVAL, "bad nli node coordinate");
                pg.push_back(pack3(nd[0], nd[1], nd[2]) | (nd[3] << 24));
                pg.push_back(nc);
                for (int k = 0; k < nc; ++k) {
                    const int mask = nd[5 + 3 * k], kind = nd[6 + 3 * k], tgt = nd[7 + 3 * k];
                    int t = tgt;
                    if (kind == 1) {
                        if (tgt < 0 || tgt > m->P)
                            return fail(KMCG_EINVAL, "bad nli leaf value");
                    } else if (kind == 0) {
                        // child: must come later in this tree (so the walk ends)
                        auto it = std::find(old_pos.begin(), old_pos.end(), tgt);
                        if (it == old_pos.end() || *it <= old_pos[q])
                            return fail(KMCG_EINVAL, "bad nli child offset");
                        t = new_pos[it - old_pos.begin()];
                    } else {
                        return fail(KMCG_EINVAL, "bad nli case kind");
                    }
                    pg.insert(pg.end(), {mask, kind, t});
                }
            }
        }
        for (size_t si = 0; si < segs.size(); ++si)
            for (size_t w = 0; w < segs[si].writes.size(); w += 6) {
                const int n = segs[si].writes[w + 3], nw = segs[si].writes[w + 5];
                if (n < 1 || n > m->spuck || !m->species_ok(nw))
                    return fail(KMCG_EINVAL, "bad lattice write");
            }
        // largest offset component any segment reaches from its base cell
        {
            int mnode = 0;
            for (const auto &src : nli_nodes) {
                size_t pos = 0;
                while (pos + 5 <= src.size()) {
                    const int nc = src[pos + 4];
                    if (nc == -1) {
                        pos += 6;
                        continue;
                    }
                    for (int q = 0; q < 3; ++q)
                        mnode = std::max(mnode, std::abs(src[pos + q]));
                    pos += 5 + 3 * (size_t)nc;
                }
            }
            int mo = 0;
            for (const auto &sg : segs) {
                for (size_t w = 0; w + 6 <= sg.writes.size(); w += 6)
                    for (int q = 0; q < 3; ++q)
                        mo = std::max(mo, std::abs(sg.writes[w + q]));
                for (const auto &op : sg.ops) {
                    const int od = std::max(std::abs(op.dx), std::max(std::abs(op.dy), std::abs(op.dz)));
                    const bool nli = op.kind == OP_DEL_NLI || op.kind == OP_ADD_NLI;
                    mo = std::max(mo, nli ? od + mnode : od);
                    for (size_t c = 0; c + 5 <= op.conds.size(); c += 5)
                        for (int q = 0; q < 3; ++q)
                            mo = std::max(mo, std::abs(op.conds[c + q]));
                }
            }
            m->moff = mo;
        }
        // nli trees over few coordinates become lookup tables: the case
        // tree is walked once here for every species combination of its
        // coordinates (first matching case, as on the device), so the device
        // reads the cluster at once and looks the variant up
        m->off_nlitab = (int)pg.size();
        pg.resize(pg.size() + nli_nodes.size(), -1);
        for (size_t g = 0; g < nli_nodes.size(); ++g) {
            const std::vector<int> &src = nli_nodes[g];
            std::vector<std::vector<int>> coords;    // dx, dy, dz, n
            std::vector<int> node_at;
            size_t pos = 0;
            while (pos + 5 <= src.size()) {
                node_at.push_back((int)pos);
                const int nc = src[pos + 4];
                if (nc == -1) {
                    pos += 6;
                    continue;
                }
                std::vector<int> c = {src[pos], src[pos + 1], src[pos + 2], src[pos + 3]};
                if (std::find(coords.begin(), coords.end(), c) == coords.end())
                    coords.push_back(c);
                pos += 5 + 3 * (size_t)nc;
            }
            // a null site (multi-layer) is one more value a coordinate can hold
            const int K = (int)coords.size(), S = m->nullsp >= 0 ? m->S + 1 : m->S;
            long long entries = 1;
            for (int k = 0; k < K && entries <= 4096; ++k)
                entries *= S;
            if (K < 1 || K > 8 || entries > 4096)
                continue;
            std::vector<unsigned char> tab((size_t)entries, 0);
            for (long long e = 0; e < entries; ++e) {
                std::vector<int> sp(K);
                long long r = e;
                for (int k = 0; k < K; ++k) {
                    sp[k] = (int)(r % S);
                    r /= S;
                }
                size_t at = 0;
                int val = 0;
                for (int guard = 0; guard < 4096; ++guard) {
                    const int *nd = &src[at];
                    if (nd[4] == -1) {
                        val = nd[5];
                        break;
                    }
                    std::vector<int> c = {nd[0], nd[1], nd[2], nd[3]};
                    const int q = sp[std::find(coords.begin(), coords.end(), c) - coords.begin()];
                    int next = -1, leaf = -1;
                    for (int k = 0; k < nd[4]; ++k) {
                        if ((nd[5 + 3 * k] >> q) & 1) {
                            if (nd[6 + 3 * k] == 1)
                                leaf = nd[7 + 3 * k];
                            else
                                next = nd[7 + 3 * k];
                            break;
                        }
                    }
                    if (leaf >= 0) {
                        val = leaf;
                        break;
                    }
                    if (next < 0) {
                        val = 0;
                        break;
                    }
                    at = (size_t)next;
                }
                if (val < 0 || val > 255)
                    return fail(KMCG_EUNSUPPORTED, "nli value does not fit a byte");
                tab[(size_t)e] = (unsigned char)val;
            }
            pg[m->off_nlitab + g] = (int)pg.size();
            pg.push_back(K | (S << 8));
            for (const auto &c : coords)
                pg.push_back(pack3(c[0], c[1], c[2]) | (c[3] << 24));
            for (size_t e = 0; e < tab.size(); e += 4) {
                unsigned w = 0;
                for (size_t q = 0; q < 4 && e + q < tab.size(); ++q)
                    w |= (unsigned)tab[e + q] << (8 * q);
                pg.push_back((int)w);
            }
            pg.insert(pg.end(), 8, 0);            // coordinate reads go 8 words ahead
        }
        // a lat_int model whose nli groups are all tabulated posts a group's
        // two segments in one pass (run_proc's mode bit 1)
        if (m->backend == 1 && m->merged && !nli_nodes.empty()) {
            bool all = true;
            for (size_t g = 0; g < nli_nodes.size(); ++g)
                all &= pg[m->off_nlitab + g] >= 0;
            if (all)
                m->merged |= 2;
        }
        // per process: anchor site type, its touchup addition op (local_smart)
        // or its group (lat_int), for the parallel initialisation.  A
        // local_smart process anchored on a site type that initialize_state
        // does not touch up (another layer than the initial one, blob word
        // 17) starts unregistered everywhere.
        const int off_itu = W(17);
        m->off_init = (int)pg.size();
        for (int p = 1; p <= m->P; ++p) {
            if (m->backend == 1) {
                pg.insert(pg.end(), {1, W(W(9) + p - 1), 0});
                continue;
            }
            int n_exec = W(W(9) + 3 * (p - 1)), opw = -1, cb = 0;
            if (n_exec < 1 || n_exec > m->spuck)
                return fail(KMCG_EINVAL, "bad anchor site type");
            const int seg = touchup[n_exec - 1];
            const int sb = m->off_seg + SEG_W * seg;
            for (size_t j = 0; j < segs[seg].ops.size(); ++j)
                if (segs[seg].ops[j].kind == OP_ADD_IF && segs[seg].ops[j].arg == p) {
                    opw = pg[sb + 3] + OP_W * (int)j;
                    cb = pg[sb + 4];
                }
            // bit 16: not registered by the initialisation; bit 17 (multi-
            // layer): the lists only promise registered => conditions hold
            const bool touched = off_itu == 0 || W(off_itu + n_exec - 1) != 0;
            pg.insert(pg.end(), {n_exec | (touched ? 0 : 1 << 16) | (m->nullsp >= 0 ? 1 << 17 : 0), opw, cb});
        }
        pg.insert(pg.end(), {0, 0, 0, 0});   // padding: condition words are read four at a time
        if (pg.size() * sizeof(int) > 64 * 1024)
            return fail(KMCG_EUNSUPPORTED, "op program larger than 64 KiB");
    } catch (const std::exception &e) {
        return fail(KMCG_EINVAL, std::string("malformed model blob: ") + e.what());
    }
    return KMCG_OK;
}

int ppl_of(int P) { return P <= 64 ? 1 : (P <= 128 ? 2 : 4); }

KParams make_params(kmcg_batch *b)
{
    // every byte zeroed, padding included: run_impl keys graphs on them
    KParams kp;
    memset(&kp, 0, sizeof kp);
    const kmcg_model *m = b->m;
    kp.prog = b->prog;
    kp.rep_ids = b->rep_ids;
    kp.dims = b->dims;
    kp.lat_off = b->lat_off_d;
    kp.avail_off = b->avail_off_d;
    kp.seed = b->seed;
    kp.stream = b->strm;
    kp.lat_pool = b->lat_pool;
    kp.avail_pool = b->avail_pool;
    kp.nr = b->nr;
    kp.rates = b->rates;
    kp.integ = b->integ;
    kp.pstat = b->pstat;
    kp.ktime = b->ktime;
    kp.kdt = b->kdt;
    kp.kstep = b->kstep;
    kp.rngctr = b->rngctr;
    kp.status = b->status;
    kp.lists_ok = b->lists_ok;
    kp.executed = b->executed;
    kp.out = b->out;
    kp.phase = b->phase;
    kp.P = m->P;
    kp.S = m->S;
    kp.spuck = m->spuck;
    kp.dim = m->dim;
    kp.maxR = m->maxR;
    kp.mrows = m->mrows;
    kp.moff = m->moff;
    kp.merged = m->merged;
    kp.backend = m->backend;
    kp.off_seg = m->off_seg;
    kp.off_proc = m->off_proc;
    kp.off_touch = m->off_touch;
    kp.off_nli = m->off_nli;
    kp.off_nlitab = m->off_nlitab;
    kp.prog_words = (int)m->prog.size();
    kp.prog_lds = (int)((m->prog.size() * sizeof(int) + 15) & ~(size_t)15);
    return kp;
}

// Launch shape for replicas `ids`: lattice in LDS when it fits, and as many
// replica-waves per workgroup (sharing one program copy) as LDS allows.
struct Shape {
    bool lds_lat;
    bool ov;           // the step kernel's overlapped form
    bool li;           // the step kernel's one-pass lat_int form
    int wpg, grid;
    size_t lds;
};

Shape launch_shape(kmcg_batch *b, const int *ids, int n, KParams &kp, int max_wpg)
{
    long long maxlb = 0;
    for (int i = 0; i < n; ++i)
        maxlb = std::max(maxlb, lattice_bytes(b->V[ids[i]], b->bits));
    const size_t occ = ((size_t)b->m->S * b->m->spuck + 3) & ~(size_t)3;
    Shape s;
    // lat_int with every nli group tabulated: the one-pass form; every
    // model with one process slot per lane keeps read-ahead back-indices
    // of the first replay block (bvq)
    s.li = (b->m->merged & 2) && ppl_of(b->m->P) == 1;
    const size_t bvq = ppl_of(b->m->P) == 1 ? (size_t)RBLK * 64 : 0;
    const size_t mbox = ((size_t)(b->m->mrows + 1) * 64 * ppl_of(b->m->P) + 64 * WIN + occ + bvq) * sizeof(int);
    const size_t prog = (size_t)kp.prog_lds;
    s.lds_lat = b->lds_mode == 0 ? false : (prog + (size_t)maxlb + mbox <= (size_t)b->max_lds);
    const size_t per_wave = (s.lds_lat ? (size_t)maxlb : 0) + mbox;
    s.wpg = std::max(1, std::min(max_wpg, std::min(MAX_WPG, n)));
    // automatic: local_smart runs with at most two replica-waves per SIMD
    // (four SIMDs per CU) and a prefix chain long enough to hide (more than
    // 16 processes; RuO2's 36 gain, ZGB's 7 lose); lat_int runs the
    // one-pass form with the replay's list words read ahead, which measured
    // faster than the overlap
    s.ov = b->ov_mode == 1 ||
           (b->ov_mode == -1 && b->m->backend == 0 && b->m->P > 16 && (long long)n <= 8LL * b->n_cu);
    while (s.wpg > 1 && prog + s.wpg * per_wave > (size_t)b->max_lds)
        s.wpg -= 1;
    s.grid = (n + s.wpg - 1) / s.wpg;
    s.lds = prog + s.wpg * per_wave;
    kp.lat_lds = s.lds_lat ? (int)maxlb : 0;
    kp.wave_lds = (int)per_wave;
    kp.wpg = s.wpg;
    kp.nrep = n;
    return s;
}

template <int PPL, int BITS>
hipError_t launch_t(bool steps, const Shape &s, hipStream_t st, const KParams &kp)
{
    const dim3 grid(s.grid), block(64 * s.wpg);
    if (steps && kp.msd_hop) {
        if (s.lds_lat)
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, true, true>), grid, block, s.lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, false, true>), grid, block, s.lds, st, kp);
    } else if (steps && PPL == 1 && s.ov) {
        if (s.lds_lat)
            hipLaunchKernelGGL((kmc_steps_kernel<1, BITS, true, false, true>), grid, block, s.lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_steps_kernel<1, BITS, false, false, true>), grid, block, s.lds, st, kp);
    } else if (steps && PPL == 1 && s.li) {
        // lat_int with tabulated nli groups: a group's segments in one pass
        if (s.lds_lat)
            hipLaunchKernelGGL((kmc_steps_kernel_li<1, BITS, true>), grid, block, s.lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_steps_kernel_li<1, BITS, false>), grid, block, s.lds, st, kp);
    } else if (steps) {
        if (s.lds_lat)
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, true, false>), grid, block, s.lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, false, false>), grid, block, s.lds, st, kp);
    } else {
        if (s.lds_lat)
            hipLaunchKernelGGL((kmc_misc_kernel<PPL, BITS, true>), grid, block, s.lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_misc_kernel<PPL, BITS, false>), grid, block, s.lds, st, kp);
    }
    return hipGetLastError();
}

template <int PPL>
hipError_t launch_b(int bits, bool steps, const Shape &s, hipStream_t st, const KParams &kp)
{
    if (bits == 2)
        return launch_t<PPL, 2>(steps, s, st, kp);
    return launch_t<PPL, 8>(steps, s, st, kp);
}

hipError_t launch(kmcg_batch *b, bool steps, const Shape &s, const KParams &kp)
{
    switch (ppl_of(b->m->P)) {
    case 1: return launch_b<1>(b->bits, steps, s, b->stream, kp);
    case 2: return launch_b<2>(b->bits, steps, s, b->stream, kp);
    default: return launch_b<4>(b->bits, steps, s, b->stream, kp);
    }
}

int check_replica(kmcg_batch *b, int r)
{
    if (!b || r < 0 || r >= b->n)
        return fail(KMCG_EINVAL, "replica index out of range");
    return KMCG_OK;
}

int settle(kmcg_batch *b);

// run the misc kernel for replicas [ids] synchronously
int run_misc(kmcg_batch *b, const std::vector<int> &ids, int mode, int a0, int a1, int a2, int a3)
{
    int rc = settle(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(b->rep_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    KParams kp = make_params(b);
    const Shape s = launch_shape(b, ids.data(), (int)ids.size(), kp, MAX_WPG);
    kp.mode = mode;
    kp.arg0 = a0; kp.arg1 = a1; kp.arg2 = a2; kp.arg3 = a3;
    HIPCHK(launch(b, false, s, kp));
    HIPCHK(hipStreamSynchronize(b->stream));
    return KMCG_OK;
}

// Wait for the batch's stream and finish the bookkeeping of a run in
// flight (its device time, its budgets back to the caller).
int settle(kmcg_batch *b)
{
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timed_run_pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->last_ms = ms;
        b->timed_run_pending = false;
    }
    if (b->user_budget) {
        memcpy(b->user_budget, b->h_budget, b->n_last * sizeof(long long));
        b->user_budget = nullptr;
    }
    if (b->user_samples) {
        memcpy(b->user_samples, b->h_samples, b->n_samples * sizeof(double));
        b->user_samples = nullptr;
    }
    if (b->user_srec) {
        memcpy(b->user_srec, b->h_srec, b->n_srec * sizeof(double));
        b->user_srec = nullptr;
    }
    return KMCG_OK;
}

template <class T>
int copy_out(kmcg_batch *b, T *dst, const T *src, size_t count)
{
    HIPCHK(hipSetDevice(b->device));
    int rc = settle(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return KMCG_OK;
}

template <class T>
int copy_in(kmcg_batch *b, T *dst, const T *src, size_t count)
{
    HIPCHK(hipSetDevice(b->device));
    int rc = settle(b);
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return KMCG_OK;
}

constexpr size_t MAX_GRAPHS = 8;

void drop_graphs(kmcg_batch *b)
{
    for (auto &c : b->graphs) {
        (void)hipGraphExecDestroy(c.exec);
        (void)hipGraphDestroy(c.graph);
    }
    b->graphs.clear();
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *kmcg_last_error(void) { return g_err.c_str(); }

int kmcg_device_count(int32_t *count)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess)
        c = 0;
    *count = c;
    return KMCG_OK;
}

int kmcg_model_load(const int32_t *blob, int64_t nwords, kmcg_model **out)
{
    if (!blob || !out || nwords < 32)
        return fail(KMCG_EINVAL, "null or short model blob");
    if (blob[0] != 0x4B4D4347 || blob[1] != 1 || blob[16] != nwords)
        return fail(KMCG_EINVAL, "not a version-1 model blob");
    kmcg_model *m = new kmcg_model();
    m->blob.assign(blob, blob + nwords);
    const int rc = build_prog(m);
    if (rc != KMCG_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return KMCG_OK;
}

void kmcg_model_free(kmcg_model *m) { delete m; }

int kmcg_model_info(const kmcg_model *m, int32_t *info5)
{
    if (!m || !info5)
        return fail(KMCG_EINVAL, "null argument");
    info5[0] = m->P; info5[1] = m->S; info5[2] = m->spuck; info5[3] = m->backend; info5[4] = m->dim;
    return KMCG_OK;
}

int kmcg_model_forms(const kmcg_model *m, int32_t *li)
{
    if (!m || !li)
        return fail(KMCG_EINVAL, "null argument");
    *li = ((m->merged & 2) && ppl_of(m->P) == 1) ? 1 : 0;
    return KMCG_OK;
}

static int sizes_ok(const kmcg_model *m, int32_t n, const int32_t *sizes)
{
    if (!m || n < 1 || !sizes)
        return fail(KMCG_EINVAL, "need a model, n >= 1 and sizes");
    for (int i = 0; i < n; ++i) {
        int X = sizes[3 * i], Y = m->dim >= 2 ? sizes[3 * i + 1] : 1, Z = m->dim >= 3 ? sizes[3 * i + 2] : 1;
        if (X < 1 || Y < 1 || Z < 1)
            return fail(KMCG_EINVAL, "system size must be positive");
        const int lim = std::min(X, std::min(m->dim >= 2 ? Y : X, m->dim >= 3 ? Z : X));
        if (lim < m->maxoff)
            return fail(KMCG_EINVAL, "system size below the model's minimum of "
                                         + std::to_string(m->maxoff) + " cells per dimension");
        if ((long long)X * Y * Z * m->spuck >= (1LL << 24))
            return fail(KMCG_EUNSUPPORTED, "more than 2^24 sites per replica");
        // the replay addresses one replica's avail tables with 32-bit byte offsets
        if (8LL * m->P * ((long long)X * Y * Z * m->spuck + 1) >= (1LL << 32))
            return fail(KMCG_EUNSUPPORTED, "avail tables of one replica exceed 4 GiB");
    }
    return KMCG_OK;
}

int kmcg_batch_footprint(const kmcg_model *m, int32_t n, const int32_t *sizes, int64_t *bytes)
{
    int rc = sizes_ok(m, n, sizes);
    if (rc)
        return rc;
    const int bits = m->lattice_bits();
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
        long long X = sizes[3 * i], Y = m->dim >= 2 ? sizes[3 * i + 1] : 1, Z = m->dim >= 3 ? sizes[3 * i + 2] : 1;
        long long V = X * Y * Z * m->spuck;
        tot += lattice_bytes((int)V, bits) + 2LL * m->P * (V + 1) * 4;
        tot += (long long)m->P * (4 + 8 + 8 + 8) + 64;
    }
    *bytes = tot;
    return KMCG_OK;
}

void kmcg_batch_free(kmcg_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    void *ptrs[] = {b->prog, b->dims, b->lat_off_d, b->avail_off_d, b->seed, b->strm, b->lat_pool,
                    b->avail_pool, b->nr, b->rates, b->integ, b->pstat, b->ktime, b->kdt, b->kstep,
                    b->rngctr, b->status, b->executed, b->budget, b->rep_ids, b->out, b->occ,
                    b->samples, b->msd_hop, b->msd_d, b->msd_id, b->msd_site, b->msd_disp,
                    b->msd_off_d, b->msd_ev, b->msd_evn, b->lists_ok, b->acf_pa, b->acf_po, b->acf_buf,
                    b->acf_pp, b->acf_bin, b->acf_cnt, b->acf_boff, b->acf_d, b->acf_i, b->phase};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (b->h_ids)
        (void)hipHostFree(b->h_ids);
    if (b->h_budget)
        (void)hipHostFree(b->h_budget);
    if (b->h_samples)
        (void)hipHostFree(b->h_samples);
    if (b->h_srec)
        (void)hipHostFree(b->h_srec);
    if (b->srec)
        (void)hipFree(b->srec);
    drop_graphs(b);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

int kmcg_batch_create(const kmcg_model *m, int32_t n, const int32_t *sizes, int32_t device, kmcg_batch **out)
{
    int rc = sizes_ok(m, n, sizes);
    if (rc)
        return rc;
    if (!out)
        return fail(KMCG_EINVAL, "null output");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(KMCG_EINVAL, "no such GPU device");
    kmcg_batch *b = new kmcg_batch();
    b->m = m;
    b->device = device;
    b->n = n;
    b->bits = m->lattice_bits();
    b->X.resize(n); b->Y.resize(n); b->Z.resize(n); b->V.resize(n);
    b->lat_off.resize(n); b->avail_off.resize(n);
    std::vector<int4> dims(n);
    for (int i = 0; i < n; ++i) {
        b->X[i] = sizes[3 * i];
        b->Y[i] = m->dim >= 2 ? sizes[3 * i + 1] : 1;
        b->Z[i] = m->dim >= 3 ? sizes[3 * i + 2] : 1;
        b->V[i] = b->X[i] * b->Y[i] * b->Z[i] * m->spuck;
        dims[i] = make_int4(b->X[i], b->Y[i], b->Z[i], b->V[i]);
        b->lat_off[i] = b->lat_total;
        b->lat_total += lattice_bytes(b->V[i], b->bits);
        b->avail_off[i] = b->avail_total;
        // 2 x P x (V+1) words, rounded to 64 words so replicas start on 256 B
        b->avail_total += ((2LL * m->P * (b->V[i] + 1)) + 63) & ~63LL;
    }
    auto bail = [&](int code) {
        kmcg_batch_free(b);
        return code;
    };
#define BALLOC(ptr, count)                                                             \
    do {                                                                               \
        if (hipMalloc((void **)&(ptr), std::max<size_t>(1, (size_t)(count)) * sizeof(*(ptr))) != hipSuccess) \
            return bail(fail(KMCG_ENOMEM, "device allocation failed for " #ptr));      \
    } while (0)
    if (hipSetDevice(device) != hipSuccess)
        return bail(fail(KMCG_EHIP, "hipSetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        b->max_lds = (int)prop.sharedMemPerBlock;
        b->n_cu = std::max(1, prop.multiProcessorCount);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(KMCG_EHIP, "stream creation failed"));
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return bail(fail(KMCG_EHIP, "event creation failed"));
    const long long P = m->P;
    BALLOC(b->prog, m->prog.size());
    BALLOC(b->dims, n);
    BALLOC(b->lat_off_d, n);
    BALLOC(b->avail_off_d, n);
    BALLOC(b->seed, n);
    BALLOC(b->strm, n);
    BALLOC(b->lat_pool, b->lat_total);
    BALLOC(b->avail_pool, b->avail_total);
    BALLOC(b->nr, n * P);
    BALLOC(b->rates, n * P);
    BALLOC(b->integ, n * P);
    BALLOC(b->pstat, n * P);
    BALLOC(b->ktime, n);
    BALLOC(b->kdt, n);
    BALLOC(b->kstep, n);
    BALLOC(b->rngctr, n);
    BALLOC(b->status, n);
    BALLOC(b->lists_ok, n);
    BALLOC(b->executed, n);
    BALLOC(b->budget, n);
    BALLOC(b->rep_ids, n);
    BALLOC(b->out, 16);
#if KMCG_PHASES
    BALLOC(b->phase, (size_t)n * PHASE_W);
#endif
    BALLOC(b->occ, (size_t)n * m->S * m->spuck);
    BALLOC(b->samples, (size_t)n * (4 + m->S * m->spuck + 2 * P));
    if (hipHostMalloc((void **)&b->h_ids, n * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_budget, n * sizeof(long long), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_samples, (size_t)n * (4 + m->S * m->spuck + 2 * P) * sizeof(double),
                      hipHostMallocDefault) != hipSuccess)
        return bail(fail(KMCG_ENOMEM, "pinned host allocation failed"));
#undef BALLOC
    auto H2D = [&](void *d, const void *h, size_t bytes) {
        return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, b->stream) == hipSuccess;
    };
    if (!H2D(b->prog, m->prog.data(), m->prog.size() * sizeof(int)) ||
        !H2D(b->dims, dims.data(), n * sizeof(int4)) ||
        !H2D(b->lat_off_d, b->lat_off.data(), n * sizeof(long long)) ||
        !H2D(b->avail_off_d, b->avail_off.data(), n * sizeof(long long)))
        return bail(fail(KMCG_EHIP, "copy to device failed"));
    if (hipMemsetAsync(b->rates, 0, n * P * sizeof(double), b->stream) != hipSuccess ||
        hipMemsetAsync(b->status, 0, n * sizeof(int), b->stream) != hipSuccess ||
        hipMemsetAsync(b->lists_ok, 0, n * sizeof(int), b->stream) != hipSuccess ||
        hipMemsetAsync(b->avail_pool, 0, b->avail_total * sizeof(int), b->stream) != hipSuccess ||
        hipMemsetAsync(b->lat_pool, 0, b->lat_total, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bail(fail(KMCG_EHIP, "memset failed"));
    *out = b;
    return KMCG_OK;
}

int kmcg_set_graph(kmcg_batch *b, int32_t enable)
{
    if (!b || enable < 0 || enable > 1)
        return fail(KMCG_EINVAL, "bad argument");
    int rc = settle(b);
    if (rc)
        return rc;
    b->use_graph = enable != 0;
    if (!b->use_graph)
        drop_graphs(b);
    return KMCG_OK;
}

int kmcg_set_overlap(kmcg_batch *b, int32_t mode)
{
    if (!b || mode < -1 || mode > 1)
        return fail(KMCG_EINVAL, "bad argument");
    b->ov_mode = mode;
    return KMCG_OK;
}

int kmcg_set_launch(kmcg_batch *b, int32_t lattice_in_lds, int32_t waves_per_group)
{
    if (!b || lattice_in_lds < -1 || lattice_in_lds > 1 || waves_per_group < 0 || waves_per_group > MAX_WPG)
        return fail(KMCG_EINVAL, "bad argument");
    b->lds_mode = lattice_in_lds;
    b->max_wpg = waves_per_group == 0 ? MAX_WPG : waves_per_group;
    return KMCG_OK;
}

static int pack_config(kmcg_batch *b, int r, const int8_t *config, std::vector<unsigned char> &packed)
{
    const int V = b->V[r];
    for (int s = 0; s < V; ++s)
        if (!b->m->species_ok(config[s]))
            return fail(KMCG_EINVAL, "configuration holds an unknown species");
    packed.assign(lattice_bytes(V, b->bits), 0);
    if (b->bits == 8) {
        for (int s = 0; s < V; ++s)
            packed[s] = (unsigned char)config[s];
    } else {
        for (int s = 0; s < V; ++s)
            packed[(s * 2) >> 3] |= (unsigned char)(((config[s] + 1) & 3) << ((s * 2) & 7));
    }
    return KMCG_OK;
}

namespace {
// mark replicas [r0, r0+n) as having lists that do (1) or may not (0) match
// their lattices
int set_lists_ok(kmcg_batch *b, int r0, int n, int v)
{
    std::vector<int> f(n, v);
    return copy_in(b, b->lists_ok + r0, f.data(), (size_t)n);
}
} // namespace

int kmcg_init(kmcg_batch *b, int32_t r0, int32_t n, const uint32_t *seeds, const uint32_t *streams,
              const int8_t *configs)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !seeds)
        return fail(KMCG_EINVAL, "bad replica range");
    int rc0 = settle(b);
    if (rc0)
        return rc0;
    // every configuration is checked before any replica is touched, so a
    // rejected call leaves all replicas as they were
    std::vector<std::vector<unsigned char>> packed;   // alive until the stream is synchronised
    if (configs) {
        long long off = 0;
        packed.resize(n);
        for (int i = 0; i < n; ++i) {
            const int rc = pack_config(b, r0 + i, configs + off, packed[i]);
            if (rc)
                return rc;
            off += b->V[r0 + i];
        }
    }
    const long long P = b->m->P;
    std::vector<unsigned> st(n, 0u);
    if (streams)
        st.assign(streams, streams + n);
    HIPCHK(hipMemcpyAsync(b->seed + r0, seeds, n * sizeof(unsigned), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->strm + r0, st.data(), n * sizeof(unsigned), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemsetAsync(b->nr + r0 * P, 0, n * P * sizeof(int), b->stream));
    HIPCHK(hipMemsetAsync(b->integ + r0 * P, 0, n * P * sizeof(double), b->stream));
    HIPCHK(hipMemsetAsync(b->pstat + r0 * P, 0, n * P * sizeof(long long), b->stream));
    HIPCHK(hipMemsetAsync(b->ktime + r0, 0, n * sizeof(double), b->stream));
    HIPCHK(hipMemsetAsync(b->kdt + r0, 0, n * sizeof(double), b->stream));
    HIPCHK(hipMemsetAsync(b->kstep + r0, 0, n * sizeof(long long), b->stream));
    HIPCHK(hipMemsetAsync(b->rngctr + r0, 0, n * sizeof(unsigned long long), b->stream));
    HIPCHK(hipMemsetAsync(b->status + r0, 0, n * sizeof(int), b->stream));
    const long long a0 = b->avail_off[r0];
    const long long a1 = (r0 + n < b->n) ? b->avail_off[r0 + n] : b->avail_total;
    HIPCHK(hipMemsetAsync(b->avail_pool + a0, 0, (a1 - a0) * sizeof(int), b->stream));
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i)
        ids[i] = r0 + i;
    HIPCHK(hipMemcpyAsync(b->rep_ids, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    KParams kp = make_params(b);
    if (configs) {
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(b->lat_pool + b->lat_off[r0 + i], packed[i].data(), packed[i].size(),
                                  hipMemcpyHostToDevice, b->stream));
    } else if (b->bits == 2) {
        hipLaunchKernelGGL((kmc_fill_default_kernel<2>), dim3(n), dim3(256), 0, b->stream, kp);
    } else {
        hipLaunchKernelGGL((kmc_fill_default_kernel<8>), dim3(n), dim3(256), 0, b->stream, kp);
    }
    HIPCHK(hipGetLastError());
    kp.arg0 = b->m->off_init;
    const dim3 grid((unsigned)(n * P));
    if (b->bits == 2)
        hipLaunchKernelGGL((kmc_init_lists_kernel<2>), grid, dim3(256), 0, b->stream, kp);
    else
        hipLaunchKernelGGL((kmc_init_lists_kernel<8>), grid, dim3(256), 0, b->stream, kp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return set_lists_ok(b, r0, n, 1);
}

int kmcg_check_lists(kmcg_batch *b, int32_t r0, int32_t n, int32_t *bad)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !bad)
        return fail(KMCG_EINVAL, "bad replica range");
    int rc = settle(b);
    if (rc)
        return rc;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i)
        ids[i] = r0 + i;
    int *dbad = nullptr;
    if (hipMalloc((void **)&dbad, (size_t)n * sizeof(int)) != hipSuccess)
        return fail(KMCG_ENOMEM, "check buffer");
    KParams kp = make_params(b);
    kp.arg0 = b->m->off_init;
    const dim3 grid((unsigned)(n * b->m->P));
    hipError_t e = hipMemcpyAsync(b->rep_ids, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(dbad, 0, (size_t)n * sizeof(int), b->stream);
    if (e == hipSuccess) {
        if (b->bits == 2)
            hipLaunchKernelGGL((kmc_check_lists_kernel<2>), grid, dim3(256), 0, b->stream, kp, dbad);
        else
            hipLaunchKernelGGL((kmc_check_lists_kernel<8>), grid, dim3(256), 0, b->stream, kp, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(bad, dbad, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    (void)hipFree(dbad);
    HIPCHK(e);
    return KMCG_OK;
}

int kmcg_set_rates(kmcg_batch *b, int32_t r0, int32_t n, const double *rates)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !rates)
        return fail(KMCG_EINVAL, "bad replica range");
    const long long P = b->m->P;
    for (long long i = 0; i < n * P; ++i)
        if (!(rates[i] >= 0.0) || !std::isfinite(rates[i]))
            return fail(KMCG_EINVAL, "rate constants must be finite and >= 0");
    return copy_in(b, b->rates + r0 * P, rates, (size_t)(n * P));
}

int kmcg_get_rates(kmcg_batch *b, int32_t r, double *rates)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    return copy_out(b, rates, b->rates + (long long)r * b->m->P, b->m->P);
}

int kmcg_set_config(kmcg_batch *b, int32_t r0, int32_t n, const int8_t *config)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !config)
        return fail(KMCG_EINVAL, "bad replica range");
    // every configuration is checked before any lattice is replaced, so a
    // rejected call leaves all replicas as they were
    long long off = 0;
    std::vector<int> ids(n);
    std::vector<std::vector<unsigned char>> packed(n);
    for (int i = 0; i < n; ++i) {
        const int rc = pack_config(b, r0 + i, config + off, packed[i]);
        if (rc)
            return rc;
        off += b->V[r0 + i];
        ids[i] = r0 + i;
    }
    for (int i = 0; i < n; ++i) {
        const int r = r0 + i;
        const int rc = copy_in(b, b->lat_pool + b->lat_off[r], packed[i].data(), packed[i].size());
        if (rc)
            return rc;
    }
    int rc = run_misc(b, ids, MODE_ADJUST, 0, 0, 0, 0);
    return rc ? rc : set_lists_ok(b, r0, n, 1);
}

// track: bit 0 displacement tracking, bit 1 trajectory, bit 2 ACF bins
static int run_impl(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                    int sample_every, int max_records, double *records, int track = 0)
{
    if (!b || !replicas || n < 1 || n > b->n || steps < 0)
        return fail(KMCG_EINVAL, "bad run arguments");
    // each replica is one wave's state: a replica named twice would be
    // stepped by two waves at once
    std::vector<unsigned char> seen(b->n, 0);
    for (int i = 0; i < n; ++i) {
        if (replicas[i] < 0 || replicas[i] >= b->n)
            return fail(KMCG_EINVAL, "replica index out of range");
        if (seen[replicas[i]]++)
            return fail(KMCG_EINVAL, "replica named twice in one run");
    }
    int rc = settle(b);      // the staging buffers belong to one run at a time
    if (rc)
        return rc;
    memcpy(b->h_ids, replicas, n * sizeof(int));
    if (budget)
        memcpy(b->h_budget, budget, n * sizeof(long long));
    KParams kp = make_params(b);
    const Shape s = launch_shape(b, replicas, n, kp, b->max_wpg);
    kp.nsteps = steps;
    kp.budget = budget ? b->budget : nullptr;
    if (track & 4) {
        kp.acf_pa = b->acf_pa;
        kp.acf_po = b->acf_po;
        kp.acf_buf = b->acf_buf;
        kp.acf_pp = b->acf_pp;
        kp.acf_nt1 = b->acf_nt1;
        kp.acf_bin = b->acf_bin;
        kp.acf_cnt = b->acf_cnt;
        kp.acf_boff = b->acf_boff;
        kp.acf_d = b->acf_d;
        kp.acf_i = b->acf_i;
    }
    if (track) {
        kp.msd_hop = b->msd_hop;
        kp.msd_d = b->msd_d;
        kp.msd_id = b->msd_id;
        kp.msd_site = b->msd_site;
        kp.msd_disp = b->msd_disp;
        kp.msd_off = b->msd_off_d;
        kp.msd_ev = b->msd_ev;
        kp.msd_evn = b->msd_evn;
        kp.msd_evcap = (track & 2) ? b->msd_evcap : 0;
    }
    const int width = 4 + b->m->S * b->m->spuck + 2 * b->m->P;
    if (records) {
        const size_t need = (size_t)n * max_records * width;
        if (need > b->srec_len) {
            if (b->srec)
                HIPCHK(hipFree(b->srec));
            if (b->h_srec)
                HIPCHK(hipHostFree(b->h_srec));
            b->srec = nullptr;
            b->h_srec = nullptr;
            b->srec_len = 0;
            if (hipMalloc((void **)&b->srec, need * sizeof(double)) != hipSuccess ||
                hipHostMalloc((void **)&b->h_srec, need * sizeof(double), hipHostMallocDefault) != hipSuccess)
                return fail(KMCG_ENOMEM, "record buffers");
            b->srec_len = need;
        }
        kp.srec = b->srec;
        kp.sample_every = sample_every;
        kp.srec_cap = max_records;
        kp.srec_width = width;
    }
    const size_t cnt = records ? (size_t)n * max_records * width : 0;
    // the stream work of the run; `timed` brackets the kernel with the
    // batch's events (not inside a graph: a replay is timed as a whole)
    auto enqueue = [&](bool timed) -> hipError_t {
        hipError_t e = hipMemcpyAsync(b->rep_ids, b->h_ids, n * sizeof(int), hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess && budget)
            e = hipMemcpyAsync(b->budget, b->h_budget, n * sizeof(long long), hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess && timed)
            e = hipEventRecord(b->ev0, b->stream);
        if (e == hipSuccess)
            e = launch(b, true, s, kp);
        if (e == hipSuccess && timed)
            e = hipEventRecord(b->ev1, b->stream);
        if (e == hipSuccess && budget)
            e = hipMemcpyAsync(b->h_budget, b->budget, n * sizeof(long long), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess && records)
            e = hipMemcpyAsync(b->h_srec, b->srec, cnt * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        return e;
    };
    if (!b->use_graph) {
        HIPCHK(enqueue(true));
    } else {
        // the key: every value a captured node holds
        const int flags[7] = {n, budget ? 1 : 0, records ? 1 : 0, s.lds_lat ? 1 : 0, s.wpg, s.grid, s.ov ? 1 : 0};
        const size_t lds = s.lds;
        const void *hs = b->h_srec;
        std::vector<unsigned char> key(sizeof kp + sizeof flags + sizeof lds + sizeof hs + sizeof cnt);
        unsigned char *kq = key.data();
        memcpy(kq, &kp, sizeof kp); kq += sizeof kp;
        memcpy(kq, flags, sizeof flags); kq += sizeof flags;
        memcpy(kq, &lds, sizeof lds); kq += sizeof lds;
        memcpy(kq, &hs, sizeof hs); kq += sizeof hs;
        memcpy(kq, &cnt, sizeof cnt);
        hipGraphExec_t exec = nullptr;
        for (auto &c : b->graphs)
            if (c.key == key)
                exec = c.exec;
        if (!exec) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue(false);
            const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g)
                    (void)hipGraphDestroy(g);
                HIPCHK(e != hipSuccess ? e : e2);
            }
            const hipError_t e3 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            if (e3 != hipSuccess) {
                (void)hipGraphDestroy(g);
                HIPCHK(e3);
            }
            if (b->graphs.size() >= MAX_GRAPHS) {
                (void)hipGraphExecDestroy(b->graphs.front().exec);
                (void)hipGraphDestroy(b->graphs.front().graph);
                b->graphs.erase(b->graphs.begin());
            }
            b->graphs.push_back({std::move(key), g, exec});
        }
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        HIPCHK(hipGraphLaunch(exec, b->stream));
        HIPCHK(hipEventRecord(b->ev1, b->stream));
    }
    b->timed_run_pending = true;
    if (budget) {
        b->user_budget = budget;
        b->n_last = n;
    }
    if (records) {
        b->user_srec = records;
        b->n_srec = cnt;
    }
    return KMCG_OK;
}

int kmcg_run(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget)
{
    return run_impl(b, replicas, n, steps, budget, 0, 0, nullptr);
}

int kmcg_run_recorded(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                      int32_t sample_every, int32_t max_records, double *records)
{
    if (!b || !records || sample_every < 1 || max_records < 1 ||
        steps > (int64_t)sample_every * max_records)
        return fail(KMCG_EINVAL, "bad record arguments (steps must fit max_records intervals)");
    if (b->m->S * b->m->spuck > 8 * 32)
        return fail(KMCG_EUNSUPPORTED, "too many species x sites for occupation");
    return run_impl(b, replicas, n, steps, budget, sample_every, max_records, records);
}


int kmcg_run_sampled(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                     double *samples)
{
    if (!samples)
        return fail(KMCG_EINVAL, "null sample buffer");
    if (b && b->m->S * b->m->spuck > 8 * 32)
        return fail(KMCG_EUNSUPPORTED, "too many species x sites for occupation");
    int rc = kmcg_run(b, replicas, n, steps, budget);
    if (rc)
        return rc;
    // the run left the replica list in rep_ids; the records follow it on the
    // same stream
    const int width = 4 + b->m->S * b->m->spuck + 2 * b->m->P;
    KParams kp = make_params(b);
    if (b->bits == 2)
        hipLaunchKernelGGL((kmc_sample_kernel<2>), dim3(n), dim3(256), 0, b->stream, kp, b->samples, width);
    else
        hipLaunchKernelGGL((kmc_sample_kernel<8>), dim3(n), dim3(256), 0, b->stream, kp, b->samples, width);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->h_samples, b->samples, (size_t)n * width * sizeof(double), hipMemcpyDeviceToHost,
                          b->stream));
    b->user_samples = samples;
    b->n_samples = (size_t)n * width;
    return KMCG_OK;
}

int kmcg_query(kmcg_batch *b, int32_t *done)
{
    if (!b || !done)
        return fail(KMCG_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipErrorNotReady) {
        *done = 0;
        return KMCG_OK;
    }
    HIPCHK(e);
    *done = 1;
    return KMCG_OK;
}

int kmcg_sync(kmcg_batch *b)
{
    if (!b)
        return fail(KMCG_EINVAL, "null batch");
    return settle(b);
}

int kmcg_last_run_ms(kmcg_batch *b, float *ms)
{
    if (!b || !ms)
        return fail(KMCG_EINVAL, "null argument");
    int rc = kmcg_sync(b);
    if (rc)
        return rc;
    *ms = b->last_ms;
    return KMCG_OK;
}

int kmcg_phase_profile(kmcg_batch *b, int32_t r0, int32_t n, uint64_t *out)
{
#if KMCG_PHASES
    if (!b || !out || r0 < 0 || n < 1 || r0 + n > b->n)
        return fail(KMCG_EINVAL, "bad replica range");
    int rc = settle(b);
    if (rc)
        return rc;
    rc = copy_out(b, (unsigned long long *)out, b->phase + (size_t)r0 * PHASE_W, (size_t)n * PHASE_W);
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(b->phase + (size_t)r0 * PHASE_W, 0, (size_t)n * PHASE_W * sizeof(unsigned long long),
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return KMCG_OK;
#else
    (void)b; (void)r0; (void)n; (void)out;
    return fail(KMCG_EUNSUPPORTED, "phase profile needs the library built with KMCG_PHASES");
#endif
}

int kmcg_get_next_step(kmcg_batch *b, int32_t r, int32_t *proc, int32_t *site)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    rc = run_misc(b, {r}, MODE_NEXT_STEP, 0, 0, 0, 0);
    if (rc)
        return rc;
    int o[2];
    rc = copy_out(b, o, b->out, 2);
    if (rc)
        return rc;
    *proc = o[0];
    *site = o[1];
    return KMCG_OK;
}

int kmcg_run_proc_nr(kmcg_batch *b, int32_t r, int32_t proc, int32_t site)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (proc < 1 || proc > b->m->P || site < 1 || site > b->V[r])
        return fail(KMCG_EINVAL, "process or site out of range");
    return run_misc(b, {r}, MODE_RUN_PROC, proc, site, 0, 0);
}

int kmcg_touchup(kmcg_batch *b, int32_t r, int32_t x, int32_t y, int32_t z, int32_t n)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (x < 0 || x >= b->X[r] || y < 0 || y >= b->Y[r] || z < 0 || z >= b->Z[r])
        return fail(KMCG_EINVAL, "cell out of range");
    if (b->m->backend == 0 && (n < 1 || n > b->m->spuck))
        return fail(KMCG_EINVAL, "site type out of range");
    return run_misc(b, {r}, MODE_TOUCHUP, x, y, z, b->m->backend == 0 ? n : 1);
}

int kmcg_replace_species(kmcg_batch *b, int32_t r, int32_t site, int32_t old_sp, int32_t new_sp)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (site < 1 || site > b->V[r] || !b->m->species_ok(new_sp))
        return fail(KMCG_EINVAL, "site or species out of range");
    std::vector<int8_t> lat(b->V[r]);
    rc = kmcg_get_lattice(b, r, lat.data());
    if (rc)
        return rc;
    if (lat[site - 1] != old_sp)
        return fail(KMCG_EINVAL, "replace_species: site holds another species");
    // the lists no longer follow the lattice until the next set_config
    if ((rc = set_lists_ok(b, r, 1, 0)))
        return rc;
    const long long s = site - 1;
    const long long byte = b->bits == 8 ? s : (s * 2) >> 3;
    unsigned char cur;
    rc = copy_out(b, &cur, b->lat_pool + b->lat_off[r] + byte, 1);
    if (rc)
        return rc;
    if (b->bits == 8) {
        cur = (unsigned char)new_sp;
    } else {
        const int sh = (int)((s * 2) & 7);
        cur = (unsigned char)((cur & ~(3u << sh)) | (((unsigned)(new_sp + 1) & 3u) << sh));
    }
    return copy_in(b, b->lat_pool + b->lat_off[r] + byte, &cur, 1);
}

int kmcg_get_scalars(kmcg_batch *b, int32_t r, double *time2, int64_t *step2, int32_t *status)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if ((rc = copy_out(b, time2, b->ktime + r, 1)) || (rc = copy_out(b, time2 + 1, b->kdt + r, 1)) ||
        (rc = copy_out(b, (long long *)step2, b->kstep + r, 1)) ||
        (rc = copy_out(b, (unsigned long long *)(step2 + 1), b->rngctr + r, 1)) ||
        (rc = copy_out(b, status, b->status + r, 1)))
        return rc;
    return KMCG_OK;
}

int kmcg_get_scalars_all(kmcg_batch *b, double *kmc_time, int64_t *kmc_step, int32_t *status)
{
    if (!b)
        return fail(KMCG_EINVAL, "null batch");
    int rc;
    if (kmc_time && (rc = copy_out(b, kmc_time, b->ktime, b->n)))
        return rc;
    if (kmc_step && (rc = copy_out(b, (long long *)kmc_step, b->kstep, b->n)))
        return rc;
    if (status && (rc = copy_out(b, status, b->status, b->n)))
        return rc;
    return KMCG_OK;
}

int kmcg_set_kmc_time(kmcg_batch *b, int32_t r, double t)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    return copy_in(b, b->ktime + r, &t, 1);
}

int kmcg_get_nr_of_sites(kmcg_batch *b, int32_t r, int32_t *out)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    return copy_out(b, out, b->nr + (long long)r * b->m->P, b->m->P);
}

int kmcg_get_procstat(kmcg_batch *b, int32_t r, int64_t *out)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    return copy_out(b, (long long *)out, b->pstat + (long long)r * b->m->P, b->m->P);
}

int kmcg_get_integ_rates(kmcg_batch *b, int32_t r, double *out)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    return copy_out(b, out, b->integ + (long long)r * b->m->P, b->m->P);
}

int kmcg_get_accum_rates(kmcg_batch *b, int32_t r, double *out)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    const int P = b->m->P;
    std::vector<int> nr(P);
    std::vector<double> rt(P);
    if ((rc = kmcg_get_nr_of_sites(b, r, nr.data())) || (rc = kmcg_get_rates(b, r, rt.data())))
        return rc;
    double a = (double)nr[0] * rt[0];
    out[0] = a;
    for (int i = 1; i < P; ++i) {
        a = a + (double)nr[i] * rt[i];
        out[i] = a;
    }
    return KMCG_OK;
}

int kmcg_update_integ_rate(kmcg_batch *b, int32_t r)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    const int P = b->m->P;
    std::vector<int> nr(P);
    std::vector<double> rt(P), ig(P);
    double dt;
    if ((rc = kmcg_get_nr_of_sites(b, r, nr.data())) || (rc = kmcg_get_rates(b, r, rt.data())) ||
        (rc = kmcg_get_integ_rates(b, r, ig.data())) || (rc = copy_out(b, &dt, b->kdt + r, 1)))
        return rc;
    for (int i = 0; i < P; ++i)
        ig[i] = ig[i] + (double)nr[i] * rt[i] * dt;
    return copy_in(b, b->integ + (long long)r * P, ig.data(), (size_t)P);
}

int kmcg_get_lattice(kmcg_batch *b, int32_t r, int8_t *out)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    const int V = b->V[r];
    std::vector<unsigned char> packed(lattice_bytes(V, b->bits));
    rc = copy_out(b, packed.data(), b->lat_pool + b->lat_off[r], packed.size());
    if (rc)
        return rc;
    for (int s = 0; s < V; ++s) {
        if (b->bits == 8)
            out[s] = (int8_t)packed[s];
        else
            out[s] = (int8_t)(((packed[(s * 2) >> 3] >> ((s * 2) & 7)) & 3) - 1);
    }
    return KMCG_OK;
}

int kmcg_get_avail(kmcg_batch *b, int32_t r, int32_t *sites, int32_t *back)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    const long long P = b->m->P, V1 = b->V[r] + 1;
    std::vector<int> blk(2 * P * V1);
    rc = copy_out(b, blk.data(), b->avail_pool + b->avail_off[r], blk.size());
    if (rc)
        return rc;
    for (long long p = 0; p < P; ++p) {
        std::memcpy(sites + p * (V1 - 1), blk.data() + p * V1 + 1, (V1 - 1) * sizeof(int));
        std::memcpy(back + p * (V1 - 1), blk.data() + (P + p) * V1 + 1, (V1 - 1) * sizeof(int));
    }
    return KMCG_OK;
}

// ---------------------------------------------------------------------------
// displacement tracking (base_acf.mpy:217-361, 825-891)
// ---------------------------------------------------------------------------

namespace {
// the hop table's sites must stay within one periodic image of every
// replica's lattice (site0 wraps once)
int check_hops(kmcg_batch *b, const int32_t *hops, const double *disp)
{
    const kmcg_model *m = b->m;
    const int P = m->P;
    // every hop's two sites must stay within one periodic image of any
    // replica's lattice (site0 wraps once)
    int minX = 1 << 30, minY = 1 << 30, minZ = 1 << 30;
    for (int r = 0; r < b->n; ++r) {
        minX = std::min(minX, b->X[r]);
        minY = std::min(minY, b->Y[r]);
        minZ = std::min(minZ, b->Z[r]);
    }
    for (int p = 0; p < P; ++p) {
        const int32_t *h = hops + MSD_HOP_W * p;
        if (h[0] != 0 && h[0] != 1)
            return fail(KMCG_EINVAL, "msd_init: hop flag must be 0 or 1");
        if (!h[0])
            continue;
        for (int e = 0; e < 2; ++e) {
            const int32_t *q = h + 1 + 4 * e;
            if (std::abs(q[0]) >= minX || (m->dim >= 2 && std::abs(q[1]) >= minY) ||
                (m->dim >= 3 && std::abs(q[2]) >= minZ) || (m->dim < 2 && q[1] != 0) ||
                (m->dim < 3 && q[2] != 0) || q[3] < 1 || q[3] > m->spuck)
                return fail(KMCG_EINVAL, "msd_init: hop site outside the lattice's reach");
        }
        for (int c = 0; disp && c < 3; ++c)
            if (!std::isfinite(disp[3 * p + c]))
                return fail(KMCG_EINVAL, "msd_init: displacements must be finite");
    }
    return KMCG_OK;
}

// the tracking pools of every replica, zero-filled (once per batch)
int msd_alloc(kmcg_batch *b)
{
    const int P = b->m->P;
    if (!b->msd_hop) {
        long long tot = 0;
        b->msd_off.assign(b->n, 0);
        for (int r = 0; r < b->n; ++r) {
            b->msd_off[r] = tot;
            tot += b->V[r] + 1;
        }
        b->msd_nions.assign(b->n, 0);
        if (hipMalloc((void **)&b->msd_hop, sizeof(int) * MSD_HOP_W * P) != hipSuccess ||
            hipMalloc((void **)&b->msd_d, sizeof(double) * 3 * P) != hipSuccess ||
            hipMalloc((void **)&b->msd_id, sizeof(int) * tot) != hipSuccess ||
            hipMalloc((void **)&b->msd_site, sizeof(int) * tot) != hipSuccess ||
            hipMalloc((void **)&b->msd_disp, sizeof(double) * 3 * tot) != hipSuccess ||
            hipMalloc((void **)&b->msd_off_d, sizeof(long long) * b->n) != hipSuccess ||
            hipMalloc((void **)&b->msd_evn, sizeof(long long) * b->n) != hipSuccess)
            return fail(KMCG_ENOMEM, "displacement tracking buffers");
        HIPCHK(hipMemcpy(b->msd_off_d, b->msd_off.data(), sizeof(long long) * b->n, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(b->msd_evn, 0, sizeof(long long) * b->n));
        // replicas never initialised track particle 0 only: their ids must be 0
        HIPCHK(hipMemset(b->msd_id, 0, sizeof(int) * tot));
        HIPCHK(hipMemset(b->msd_site, 0, sizeof(int) * tot));
        HIPCHK(hipMemset(b->msd_disp, 0, sizeof(double) * 3 * tot));
        HIPCHK(hipMemset(b->msd_d, 0, sizeof(double) * 3 * P));
    }
    return KMCG_OK;
}

// particles numbered 1.. over the sites holding trace_species, in site
// order (calc_nr_of_ions + the numbering of initialize_acf /
// initialize_mean_squared_displacement); displacements to zero when asked
int msd_number(kmcg_batch *b, int r0, int n, int trace_species, bool zero_disp)
{
    int rc;
    for (int r = r0; r < r0 + n; ++r) {
        const int V = b->V[r];
        std::vector<int8_t> lat(V);
        if ((rc = kmcg_get_lattice(b, r, lat.data())))
            return rc;
        std::vector<int> id(V + 1, 0), sa(V + 1, 0);
        int k = 0;
        for (int s = 1; s <= V; ++s)
            if (lat[s - 1] == trace_species) {
                ++k;
                id[s] = k;
                sa[k] = s;
            }
        b->msd_nions[r] = k;
        const long long o = b->msd_off[r];
        std::vector<double> zero(3 * (size_t)(V + 1), 0.0);
        const long long none = 0;
        if ((rc = copy_in(b, b->msd_id + o, id.data(), id.size())) ||
            (rc = copy_in(b, b->msd_site + o, sa.data(), sa.size())) ||
            (zero_disp && (rc = copy_in(b, b->msd_disp + 3 * o, zero.data(), zero.size()))) ||
            (rc = copy_in(b, b->msd_evn + r, &none, 1)))
            return rc;
    }
    return KMCG_OK;
}
} // namespace

int kmcg_msd_init(kmcg_batch *b, int32_t r0, int32_t n, int32_t trace_species, const int32_t *hops,
                  const double *disp)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !hops || !disp)
        return fail(KMCG_EINVAL, "bad msd_init arguments");
    if (trace_species < 0 || trace_species >= b->m->S)
        return fail(KMCG_EINVAL, "msd_init: unknown trace species");
    int rc = check_hops(b, hops, disp);
    if (rc || (rc = settle(b)) || (rc = msd_alloc(b)))
        return rc;
    HIPCHK(hipMemcpy(b->msd_hop, hops, sizeof(int) * MSD_HOP_W * b->m->P, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->msd_d, disp, sizeof(double) * 3 * b->m->P, hipMemcpyHostToDevice));
    return msd_number(b, r0, n, trace_species, true);
}

int kmcg_msd_trajectory(kmcg_batch *b, int64_t nr_of_steps)
{
    if (!b || nr_of_steps < 0 || nr_of_steps > (1ll << 31))
        return fail(KMCG_EINVAL, "bad trajectory length");
    if (!b->msd_hop)
        return fail(KMCG_EINVAL, "msd_trajectory before msd_init");
    int rc = settle(b);
    if (rc)
        return rc;
    if (b->msd_ev)
        HIPCHK(hipFree(b->msd_ev));
    b->msd_ev = nullptr;
    b->msd_evcap = 0;
    if (nr_of_steps > 0) {
        if (hipMalloc((void **)&b->msd_ev, sizeof(int) * 2 * (size_t)nr_of_steps * b->n) != hipSuccess)
            return fail(KMCG_ENOMEM, "trajectory buffer");
        b->msd_evcap = nr_of_steps;
    }
    HIPCHK(hipMemset(b->msd_evn, 0, sizeof(long long) * b->n));
    return KMCG_OK;
}

int kmcg_msd_zero(kmcg_batch *b, int32_t r0, int32_t n)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n)
        return fail(KMCG_EINVAL, "bad replica range");
    if (!b->msd_hop)
        return fail(KMCG_EINVAL, "msd_zero before msd_init");
    int rc = settle(b);
    if (rc)
        return rc;
    const long long o0 = b->msd_off[r0];
    const long long o1 = b->msd_off[r0 + n - 1] + b->V[r0 + n - 1] + 1;
    HIPCHK(hipMemset(b->msd_disp + 3 * o0, 0, sizeof(double) * 3 * (size_t)(o1 - o0)));
    return KMCG_OK;
}

int kmcg_msd_get(kmcg_batch *b, int32_t r, int32_t *nr_of_ions, int32_t *site_arr, int32_t *id_arr,
                 double *displacement, int64_t *n_events, int32_t *events)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (!b->msd_hop)
        return fail(KMCG_EINVAL, "msd_get before msd_init");
    const long long o = b->msd_off[r], V1 = b->V[r] + 1;
    long long ne = 0;
    if ((rc = copy_out(b, &ne, (const long long *)b->msd_evn + r, 1)))
        return rc;
    if (nr_of_ions)
        *nr_of_ions = b->msd_nions[r];
    if (n_events)
        *n_events = ne;
    if ((site_arr && (rc = copy_out(b, site_arr, (const int *)b->msd_site + o, (size_t)V1))) ||
        (id_arr && (rc = copy_out(b, id_arr, (const int *)b->msd_id + o, (size_t)V1))) ||
        (displacement && (rc = copy_out(b, displacement, (const double *)b->msd_disp + 3 * o, (size_t)(3 * V1)))))
        return rc;
    if (events && ne > 0 && b->msd_ev)
        if ((rc = copy_out(b, events, (const int *)b->msd_ev + 2 * (size_t)r * b->msd_evcap, (size_t)(2 * ne))))
            return rc;
    return KMCG_OK;
}

int kmcg_run_displacement(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                          int32_t traj_on)
{
    if (!b || !b->msd_hop)
        return fail(KMCG_EINVAL, "run_displacement before msd_init");
    return run_impl(b, replicas, n, steps, budget, 0, 0, nullptr, traj_on ? 3 : 1);
}

// ---------------------------------------------------------------------------
// autocorrelation bins (base_acf.mpy:382-552, 1053-1201)
// ---------------------------------------------------------------------------

namespace {
// update_property_and_buffer_acf and sum(buffer_acf) of replica r, on the host
int acf_refresh_host(kmcg_batch *b, int r)
{
    const long long o = b->msd_off[r], V1 = b->V[r] + 1;
    std::vector<int> sa(V1), pa(V1), po(V1);
    std::vector<double> buf(V1), pp((size_t)b->acf_nt1 * b->acf_nt1);
    int rc;
    if ((rc = copy_out(b, sa.data(), (const int *)b->msd_site + o, (size_t)V1)) ||
        (rc = copy_out(b, pa.data(), (const int *)b->acf_pa + o, (size_t)V1)) ||
        (rc = copy_out(b, po.data(), (const int *)b->acf_po + o, (size_t)V1)) ||
        (rc = copy_out(b, buf.data(), (const double *)b->acf_buf + o, (size_t)V1)) ||
        (rc = copy_out(b, pp.data(), (const double *)b->acf_pp, pp.size())))
        return rc;
    double sum = buf[0];
    for (long long i = 1; i < V1; ++i) {
        po[i] = pa[sa[i]];
        buf[i] = pp[(size_t)po[i] * b->acf_nt1 + po[i]];
        sum = sum + buf[i];
    }
    if ((rc = copy_in(b, b->acf_po + o, po.data(), (size_t)V1)) ||
        (rc = copy_in(b, b->acf_buf + o, buf.data(), (size_t)V1)) ||
        (rc = copy_in(b, b->acf_d + 8 * (long long)r, &sum, 1)))
        return rc;
    return KMCG_OK;
}
} // namespace

int kmcg_acf_alloc(kmcg_batch *b, int32_t nr_of_types, double t_bin, double t_f, int32_t safety_factor,
                   int32_t extending_factor)
{
    if (!b || nr_of_types < 1 || nr_of_types > 4096 || !(t_bin > 0.0) || !(t_f > 0.0) ||
        !std::isfinite(t_bin) || !std::isfinite(t_f) || safety_factor < 0 || extending_factor < 0)
        return fail(KMCG_EINVAL, "bad acf_alloc arguments");
    const int safety = safety_factor == 0 ? 2 : safety_factor;
    const int extend = extending_factor == 0 ? 2 : extending_factor;
    const double tf = t_f * safety;
    const double nbd = tf / t_bin + 1;
    if (!(nbd < 1e7))
        return fail(KMCG_EINVAL, "acf_alloc: too many bins");
    const int nb = (int)nbd;
    const long long ext = (long long)nb * extend;
    if (ext * b->n > (1ll << 28))
        return fail(KMCG_EINVAL, "acf_alloc: too many bins for this batch");
    int rc = settle(b);
    if (rc || (rc = msd_alloc(b)))
        return rc;
    void *old[] = {b->acf_pa, b->acf_po, b->acf_buf, b->acf_pp, b->acf_bin, b->acf_cnt, b->acf_boff, b->acf_d,
                   b->acf_i};
    for (void *p : old)
        if (p)
            HIPCHK(hipFree(p));
    long long tot = 0;
    for (int r = 0; r < b->n; ++r)
        tot += b->V[r] + 1;
    const int nt1 = nr_of_types + 1;
    if (hipMalloc((void **)&b->acf_pa, sizeof(int) * tot) != hipSuccess ||
        hipMalloc((void **)&b->acf_po, sizeof(int) * tot) != hipSuccess ||
        hipMalloc((void **)&b->acf_buf, sizeof(double) * tot) != hipSuccess ||
        hipMalloc((void **)&b->acf_pp, sizeof(double) * nt1 * nt1) != hipSuccess ||
        hipMalloc((void **)&b->acf_bin, sizeof(double) * ext * b->n) != hipSuccess ||
        hipMalloc((void **)&b->acf_cnt, sizeof(int) * ext * b->n) != hipSuccess ||
        hipMalloc((void **)&b->acf_boff, sizeof(long long) * b->n) != hipSuccess ||
        hipMalloc((void **)&b->acf_d, sizeof(double) * 8 * b->n) != hipSuccess ||
        hipMalloc((void **)&b->acf_i, sizeof(int) * 8 * b->n) != hipSuccess)
        return fail(KMCG_ENOMEM, "acf buffers");
    b->acf_nt1 = nt1;
    b->acf_ext = (int)ext;
    HIPCHK(hipMemset(b->acf_pa, 0, sizeof(int) * tot));
    HIPCHK(hipMemset(b->acf_po, 0, sizeof(int) * tot));
    HIPCHK(hipMemset(b->acf_buf, 0, sizeof(double) * tot));
    HIPCHK(hipMemset(b->acf_pp, 0, sizeof(double) * nt1 * nt1));
    HIPCHK(hipMemset(b->acf_bin, 0, sizeof(double) * ext * b->n));
    HIPCHK(hipMemset(b->acf_cnt, 0, sizeof(int) * ext * b->n));
    std::vector<long long> boff(b->n);
    std::vector<double> d(8 * (size_t)b->n, 0.0), kt(b->n);
    std::vector<int> iv(8 * (size_t)b->n, 0);
    HIPCHK(hipMemcpy(kt.data(), b->ktime, sizeof(double) * b->n, hipMemcpyDeviceToHost));
    for (int r = 0; r < b->n; ++r) {
        boff[r] = (long long)r * ext;
        d[8 * r + 1] = kt[r];                 // t_o
        d[8 * r + 3] = t_bin;
        iv[8 * r + 1] = nb;
        iv[8 * r + 2] = (int)ext;
        iv[8 * r + 3] = b->msd_nions[r];
        iv[8 * r + 4] = b->msd_nions[r];
    }
    HIPCHK(hipMemcpy(b->acf_boff, boff.data(), sizeof(long long) * b->n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->acf_d, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->acf_i, iv.data(), sizeof(int) * iv.size(), hipMemcpyHostToDevice));
    return KMCG_OK;
}

int kmcg_acf_set_properties(kmcg_batch *b, int32_t r0, int32_t n, const int32_t *property_acf,
                            const double *product_property)
{
    if (!b || !b->acf_pa)
        return fail(KMCG_EINVAL, "acf_set_properties before acf_alloc");
    if (r0 < 0 || n < 0 || r0 + n > b->n || (n > 0 && !property_acf))
        return fail(KMCG_EINVAL, "bad replica range");
    const int nt1 = b->acf_nt1;
    // everything is checked before anything is copied, so a rejected call
    // changes nothing
    if (product_property) {
        // sum(buffer_acf) is kept as a running sum: exact only for integers
        for (int i = 0; i < nt1 * nt1; ++i) {
            const double v = product_property[i];
            if (!(std::fabs(v) <= 1048576.0) || v != std::floor(v))
                return fail(KMCG_EUNSUPPORTED, "acf: product properties must be integers of magnitude <= 2^20");
        }
    }
    long long off = 0;
    for (int r = r0; r < r0 + n; ++r) {
        const long long V1 = b->V[r] + 1;
        for (long long i = 0; i < V1; ++i)
            if (property_acf[off + i] < 0 || property_acf[off + i] >= nt1)
                return fail(KMCG_EINVAL, "acf: property type out of range");
        off += V1;
    }
    if (product_property) {
        const int rc = copy_in(b, b->acf_pp, product_property, (size_t)nt1 * nt1);
        if (rc)
            return rc;
    }
    off = 0;
    for (int r = r0; r < r0 + n; ++r) {
        const long long V1 = b->V[r] + 1;
        const int rc = copy_in(b, b->acf_pa + b->msd_off[r], property_acf + off, (size_t)V1);
        if (rc)
            return rc;
        off += V1;
    }
    return KMCG_OK;
}

int kmcg_acf_init(kmcg_batch *b, int32_t r0, int32_t n, int32_t trace_species, const int32_t *hops)
{
    if (!b || !b->acf_pa)
        return fail(KMCG_EINVAL, "acf_init before acf_alloc");
    if (r0 < 0 || n < 1 || r0 + n > b->n)
        return fail(KMCG_EINVAL, "bad replica range");
    if (!hops || trace_species < 0 || trace_species >= b->m->S)
        return fail(KMCG_EINVAL, "bad acf_init arguments");
    // initialize_acf: the particle numbering (displacements untouched), then
    // update_property_and_buffer_acf
    int rc = check_hops(b, hops, nullptr);
    if (rc || (rc = settle(b)))
        return rc;
    HIPCHK(hipMemcpy(b->msd_hop, hops, sizeof(int) * MSD_HOP_W * b->m->P, hipMemcpyHostToDevice));
    if ((rc = msd_number(b, r0, n, trace_species, false)))
        return rc;
    for (int r = r0; r < r0 + n; ++r) {
        const int nions = b->msd_nions[r];
        int two[2] = {nions, nions};
        if ((rc = copy_in(b, b->acf_i + 8 * (long long)r + 3, two, 2)) || (rc = acf_refresh_host(b, r)))
            return rc;
    }
    return KMCG_OK;
}

int kmcg_acf_zero(kmcg_batch *b, int32_t r0, int32_t n)
{
    if (!b || !b->acf_pa)
        return fail(KMCG_EINVAL, "acf_zero before acf_alloc");
    if (r0 < 0 || n < 1 || r0 + n > b->n)
        return fail(KMCG_EINVAL, "bad replica range");
    int rc = settle(b);
    if (rc)
        return rc;
    for (int r = r0; r < r0 + n; ++r) {
        // set_acf_to_zero: nr_of_ions_before, property/buffer, bins, t_o
        int iv[8];
        double d[8], kt;
        if ((rc = copy_out(b, iv, (const int *)b->acf_i + 8 * (long long)r, 8)) ||
            (rc = copy_out(b, d, (const double *)b->acf_d + 8 * (long long)r, 8)) ||
            (rc = copy_out(b, &kt, (const double *)b->ktime + r, 1)))
            return rc;
        iv[3] = iv[4];
        iv[0] = 0;
        d[1] = kt;
        d[2] = 0.0;
        if ((rc = copy_in(b, b->acf_i + 8 * (long long)r, iv, 8)) ||
            (rc = copy_in(b, b->acf_d + 8 * (long long)r, d, 8)) || (rc = acf_refresh_host(b, r)))
            return rc;
        HIPCHK(hipMemset(b->acf_bin + (long long)r * b->acf_ext, 0, sizeof(double) * b->acf_ext));
        HIPCHK(hipMemset(b->acf_cnt + (long long)r * b->acf_ext, 0, sizeof(int) * b->acf_ext));
    }
    return KMCG_OK;
}

int kmcg_run_acf(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget,
                 int32_t traj_on)
{
    if (!b || !b->acf_pa)
        return fail(KMCG_EINVAL, "run_acf before acf_alloc");
    return run_impl(b, replicas, n, steps, budget, 0, 0, nullptr, traj_on ? 7 : 5);
}

int kmcg_acf_get(kmcg_batch *b, int32_t r, double *config_bin, int32_t *counter_write_in_bin, double *d4,
                 int32_t *i5, double *buffer_acf, int32_t *property_o)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (!b->acf_pa)
        return fail(KMCG_EINVAL, "acf_get before acf_alloc");
    const long long o = b->msd_off[r], V1 = b->V[r] + 1, e = b->acf_ext;
    if ((config_bin && (rc = copy_out(b, config_bin, (const double *)b->acf_bin + r * e, (size_t)e))) ||
        (counter_write_in_bin && (rc = copy_out(b, counter_write_in_bin, (const int *)b->acf_cnt + r * e, (size_t)e))) ||
        (d4 && (rc = copy_out(b, d4, (const double *)b->acf_d + 8 * (long long)r, 4))) ||
        (i5 && (rc = copy_out(b, i5, (const int *)b->acf_i + 8 * (long long)r, 5))) ||
        (buffer_acf && (rc = copy_out(b, buffer_acf, (const double *)b->acf_buf + o, (size_t)V1))) ||
        (property_o && (rc = copy_out(b, property_o, (const int *)b->acf_po + o, (size_t)V1))))
        return rc;
    return KMCG_OK;
}

int kmcg_set_state(kmcg_batch *b, int32_t r, const double *time2, const int64_t *step2,
                   const int8_t *lattice, const int32_t *nr_of_sites, const int64_t *procstat,
                   const double *integ_rates, const int32_t *sites, const int32_t *back)
{
    int rc = check_replica(b, r);
    if (rc)
        return rc;
    if (!time2 || !step2 || !lattice || !nr_of_sites || !procstat || !sites || !back)
        return fail(KMCG_EINVAL, "set_state: null argument");
    if (!std::isfinite(time2[0]) || !std::isfinite(time2[1]) || step2[0] < 0)
        return fail(KMCG_EINVAL, "set_state: kmc_time, kmc_time_step must be finite, kmc_step >= 0");
    const long long P = b->m->P, V = b->V[r], V1 = V + 1;
    // every list index and site the kernels will follow is checked here:
    // an inconsistent list would send add_proc/del_proc out of the row
    std::vector<int> blk(2 * P * V1, 0);
    for (long long p = 0; p < P; ++p) {
        const long long nr = nr_of_sites[p];
        if (nr < 0 || nr > V)
            return fail(KMCG_EINVAL, "set_state: nr_of_sites out of range");
        const int32_t *sp = sites + p * V, *bp = back + p * V;
        long long held = 0;
        for (long long i = 1; i <= V; ++i) {
            const int s = sp[i - 1];
            if (i <= nr) {
                if (s < 1 || s > V || bp[s - 1] != i)
                    return fail(KMCG_EINVAL, "set_state: avail_sites and avail_sites_back disagree");
            } else if (s != 0) {
                return fail(KMCG_EINVAL, "set_state: avail_sites holds entries beyond nr_of_sites");
            }
            held += bp[i - 1] != 0;
            blk[p * V1 + i] = s;
            blk[(P + p) * V1 + i] = bp[i - 1];
        }
        if (held != nr)     // a back entry pointing at a slot that does not point back
            return fail(KMCG_EINVAL, "set_state: avail_sites_back holds sites not in the list");
    }
    for (long long p = 0; p < P; ++p)
        if (integ_rates && !std::isfinite(integ_rates[p]))
            return fail(KMCG_EINVAL, "set_state: integ_rates must be finite");
    std::vector<unsigned char> packed;
    if ((rc = pack_config(b, r, lattice, packed)))
        return rc;
    const std::vector<double> ig = integ_rates ? std::vector<double>(integ_rates, integ_rates + P)
                                               : std::vector<double>(P, 0.0);
    const int running = 0;
    if ((rc = copy_in(b, b->lat_pool + b->lat_off[r], packed.data(), packed.size())) ||
        (rc = copy_in(b, b->avail_pool + b->avail_off[r], blk.data(), blk.size())) ||
        (rc = copy_in(b, b->nr + r * P, (const int *)nr_of_sites, (size_t)P)) ||
        (rc = copy_in(b, b->pstat + r * P, (const long long *)procstat, (size_t)P)) ||
        (rc = copy_in(b, b->integ + r * P, ig.data(), (size_t)P)) ||
        (rc = copy_in(b, b->ktime + r, time2, 1)) || (rc = copy_in(b, b->kdt + r, time2 + 1, 1)) ||
        (rc = copy_in(b, b->kstep + r, (const long long *)step2, 1)) ||
        (rc = copy_in(b, b->rngctr + r, (const unsigned long long *)(step2 + 1), 1)) ||
        (rc = copy_in(b, b->status + r, &running, 1)))
        return rc;
    return set_lists_ok(b, r, 1, 0);      // a restored state is not checked against its lattice
}

int kmcg_get_occupation(kmcg_batch *b, int32_t r0, int32_t n, double *out)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !out)
        return fail(KMCG_EINVAL, "bad replica range");
    if (b->m->S * b->m->spuck > 8 * 32)
        return fail(KMCG_EUNSUPPORTED, "too many species x sites for occupation");
    HIPCHK(hipSetDevice(b->device));
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i)
        ids[i] = r0 + i;
    HIPCHK(hipMemcpyAsync(b->rep_ids, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    KParams kp = make_params(b);
    if (b->bits == 2)
        hipLaunchKernelGGL((kmc_occupation_kernel<2>), dim3(n), dim3(256), 0, b->stream, kp, b->occ);
    else
        hipLaunchKernelGGL((kmc_occupation_kernel<8>), dim3(n), dim3(256), 0, b->stream, kp, b->occ);
    HIPCHK(hipGetLastError());
    return copy_out(b, out, b->occ, (size_t)n * b->m->S * b->m->spuck);
}

int kmcg_get_counters(kmcg_batch *b, int32_t r0, int32_t n, int64_t *procstat, double *integ)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n)
        return fail(KMCG_EINVAL, "bad replica range");
    const long long P = b->m->P;
    int rc;
    if (procstat && (rc = copy_out(b, (long long *)procstat, b->pstat + r0 * P, (size_t)(n * P))))
        return rc;
    if (integ && (rc = copy_out(b, integ, b->integ + r0 * P, (size_t)(n * P))))
        return rc;
    return KMCG_OK;
}

int kmcg_sample(kmcg_batch *b, const int32_t *replicas, int32_t n, double *out)
{
    if (!b || !replicas || n < 1 || n > b->n || !out)
        return fail(KMCG_EINVAL, "bad sample arguments");
    for (int i = 0; i < n; ++i)
        if (replicas[i] < 0 || replicas[i] >= b->n)
            return fail(KMCG_EINVAL, "replica index out of range");
    if (b->m->S * b->m->spuck > 8 * 32)
        return fail(KMCG_EUNSUPPORTED, "too many species x sites for occupation");
    HIPCHK(hipSetDevice(b->device));
    const int width = 4 + b->m->S * b->m->spuck + 2 * b->m->P;
    HIPCHK(hipMemcpyAsync(b->rep_ids, replicas, n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    KParams kp = make_params(b);
    if (b->bits == 2)
        hipLaunchKernelGGL((kmc_sample_kernel<2>), dim3(n), dim3(256), 0, b->stream, kp, b->samples, width);
    else
        hipLaunchKernelGGL((kmc_sample_kernel<8>), dim3(n), dim3(256), 0, b->stream, kp, b->samples, width);
    HIPCHK(hipGetLastError());
    return copy_out(b, out, b->samples, (size_t)n * width);
}

} // extern "C"

// the otf backend (same translation unit: error helpers, uniform stream)
#include "kmcg_otf.hip"
