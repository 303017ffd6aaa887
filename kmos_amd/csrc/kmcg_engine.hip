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

using namespace kmcg;

// ===========================================================================
// device code
// ===========================================================================
namespace {

constexpr int MODE_INIT = 0;       // touchups in k, j, i order (initialize_state)
constexpr int MODE_ADJUST = 1;     // touchups with x outermost (_adjust_database)
constexpr int MODE_TOUCHUP = 2;    // one touchup at (arg0..arg3)
constexpr int MODE_RUN_PROC = 3;   // run_proc_nr(arg0, arg1)
constexpr int MODE_NEXT_STEP = 4;  // get_next_kmc_step -> out[0..1]

template <int PPL, int BITS>
struct Rep {
    const int *prog;
    Dims d;
    Lattice<BITS> lat;
    int *mbox;
    int *avail;        // this replica's avail block: sites[P][V+1] then back[P][V+1]
    long long V1;
    int P, lane, mstride;
    int nr[PPL];
    double rate[PPL];
    double integ[PPL];
    long long pstat[PPL];

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
template <int BITS>
__device__ int nli_eval(const int *prog, int pos, const Lattice<BITS> &lat, const Dims &d,
                        int cx, int cy, int cz)
{
    for (;;) {
        const int *nd = prog + pos;
        const int nc = nd[4];
        if (nc == -1)
            return nd[5];
        const int sp = lat.get(site0(d, cx, cy, cz, nd[0], nd[1], nd[2], nd[3]));
        int next = -1;
        for (int k = 0; k < nc; ++k) {
            const int *e = nd + 5 + 3 * k;
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

// Execute one segment at base cell (bx, by, bz).  Wave-uniform call.
template <int PPL, int BITS>
__device__ int run_segment(Rep<PPL, BITS> &R, int seg, int bx, int by, int bz)
{
    const int *prog = R.prog;
    const int *sg = prog + prog[H_SEG] + SEG_W * seg;
    const int nw = sg[0], woff = sg[1], nops = sg[2], ooff = sg[3], nrounds = sg[4];

    // lattice writes with the replace_species check (base.mpy:1187-1231)
    for (int w = 0; w < nw; ++w) {
        const int *wr = prog + woff + WR_W * w;
        const int s = site0(R.d, bx, by, bz, wr[0], wr[1], wr[2], wr[3]);
        const int cur = R.lat.get(s);
        if (cur < 0 || !((wr[4] >> cur) & 1))
            return ST_MISMATCH;
        __syncthreads();
        if (R.lane == 0)
            R.lat.set(s, wr[5]);
        __syncthreads();
    }
    __syncthreads();

    // evaluate ops lane-parallel and post them to the owner's mailbox
    for (int base = 0; base < nops; base += 64) {
        const int j = base + R.lane;
        if (j < nops) {
            const int *op = prog + ooff + OP_W * j;
            const int kind = op[0], rnd = op[1], arg = op[2];
            int proc = 0, code = 0, s;
            if (kind == OP_DEL_IF || kind == OP_ADD_IF) {
                s = site0(R.d, bx, by, bz, op[3], op[4], op[5], op[6]);
                if (kind == OP_DEL_IF) {
                    proc = arg;
                    code = MB_DEL_IF;
                } else {
                    bool ok = true;
                    const int *cd = prog + op[8];
                    for (int c = 0; c < op[7]; ++c, cd += CD_W) {
                        if (R.lat.get(site0(R.d, bx, by, bz, cd[0], cd[1], cd[2], cd[3])) != cd[4]) {
                            ok = false;
                            break;
                        }
                    }
                    if (ok) {
                        proc = arg;
                        code = MB_ADD;
                    }
                }
            } else {
                const int cx = wrap1(bx + op[3], R.d.X);
                const int cy = R.d.dim >= 2 ? wrap1(by + op[4], R.d.Y) : 0;
                const int cz = R.d.dim >= 3 ? wrap1(bz + op[5], R.d.Z) : 0;
                s = site0(R.d, cx, cy, cz, 0, 0, 0, 1);
                proc = nli_eval<BITS>(prog, prog[prog[H_NLI] + arg], R.lat, R.d, cx, cy, cz);
                code = (kind == OP_DEL_NLI) ? MB_DEL_STRICT : MB_ADD;
            }
            if (proc > 0)
                R.mbox[rnd * R.mstride + proc - 1] = ((s + 1) << 2) | code;
        }
    }
    __syncthreads();

    // apply: in round r each lane executes the op posted for its process(es)
    int err = 0;
    for (int r = 0; r < nrounds; ++r) {
        int v[PPL], m[PPL], moved[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const int p = R.lane + 64 * k;
            v[k] = 0;
            if (p < R.P) {
                int *slot = R.mbox + r * R.mstride + p;
                v[k] = *slot;
                if (v[k])
                    *slot = 0;
            }
        }
        // issue the reads of every slot before any write (different
        // processes never alias)
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            m[k] = 0;
            moved[k] = 0;
            if (v[k] && (v[k] & 3) != MB_ADD) {
                const int p = R.lane + 64 * k;
                m[k] = R.back(p)[v[k] >> 2];
                moved[k] = R.sites(p)[R.nr[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (!v[k])
                continue;
            const int p = R.lane + 64 * k;
            const int site = v[k] >> 2, code = v[k] & 3;
            int *sp = R.sites(p);
            int *bp = R.back(p);
            if (code == MB_ADD) {             // base.mpy:268-301 add_proc
                const int n1 = R.nr[k] + 1;
                if (n1 > R.d.V) {             // only a double add can get here
                    err |= 2;
                    continue;
                }
                sp[n1] = site;
                bp[site] = n1;
                R.nr[k] = n1;
            } else if (m[k] == 0) {           // not registered
                if (code == MB_DEL_STRICT)
                    err |= 1;
            } else {                          // base.mpy:211-265 del_proc
                const int last = R.nr[k];
                if (m[k] < last) {
                    sp[m[k]] = moved[k];
                    sp[last] = 0;
                    bp[moved[k]] = m[k];
                } else {
                    sp[m[k]] = 0;
                }
                bp[site] = 0;
                R.nr[k] = last - 1;
            }
        }
    }
    if (__ballot(err & 2))
        return ST_CORRUPT;
    return __ballot(err & 1) ? ST_DEL_ABSENT : ST_OK;
}

// run_proc_nr(proc, site): procstat is counted by the caller.
template <int PPL, int BITS>
__device__ int run_proc(Rep<PPL, BITS> &R, int proc, int site)
{
    const int s0 = site - 1;
    const int cell = s0 / R.d.spuck;
    const int x = cell % R.d.X;
    const int y = (cell / R.d.X) % R.d.Y;
    const int z = cell / (R.d.X * R.d.Y);
    const int *pp = R.prog + R.prog[H_PROG] + 2 * (proc - 1);
    const int nit = pp[0];
    const int *it = R.prog + pp[1];
    for (int i = 0; i < nit; ++i, it += IT_W) {
        const int bx = wrap1(x + it[1], R.d.X);
        const int by = R.d.dim >= 2 ? wrap1(y + it[2], R.d.Y) : 0;
        const int bz = R.d.dim >= 3 ? wrap1(z + it[3], R.d.Z) : 0;
        const int e = run_segment<PPL, BITS>(R, it[0], bx, by, bz);
        if (e)
            return e;
    }
    return ST_OK;
}

// accum_rates (base.mpy:603-623) as a strictly left-to-right FP64 sum; the
// partial sum of process p ends on lane p%64, slot p/64.  Returns the total.
template <int PPL>
__device__ __forceinline__ double accumulate(const double (&c)[PPL], double (&acc)[PPL], int P, int lane)
{
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int lim = min(64, P - 64 * k);
        acc[k] = 0.0;
        for (int l = 0; l < lim; ++l) {
            const double cl = readlane_d(c[k], l);
            a = (k == 0 && l == 0) ? cl : a + cl;
            acc[k] = (lane == l) ? a : acc[k];
        }
    }
    return a;
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

template <int PPL, int BITS, bool LDS_LAT>
__device__ __forceinline__ void rep_load(Rep<PPL, BITS> &R, const KParams &kp, int r, unsigned char *smem)
{
    R.prog = kp.prog;
    R.lane = threadIdx.x;
    R.P = kp.P;
    R.mstride = 64 * PPL;
    const int4 dd = kp.dims[r];
    R.d.X = dd.x; R.d.Y = dd.y; R.d.Z = dd.z; R.d.V = dd.w;
    R.d.spuck = kp.spuck; R.d.dim = kp.dim;
    R.V1 = (long long)dd.w + 1;
    R.avail = kp.avail_pool + kp.avail_off[r];
    unsigned char *glat = kp.lat_pool + kp.lat_off[r];
    const long long lb = lattice_bytes(dd.w, BITS);
    if (LDS_LAT) {
        const uint4 *src = (const uint4 *)glat;
        uint4 *dst = (uint4 *)smem;
        for (long long i = R.lane; i < lb / 16; i += 64)
            dst[i] = src[i];
        R.lat.p = smem;
        R.mbox = (int *)(smem + kp.lat_lds);
    } else {
        R.lat.p = glat;
        R.mbox = (int *)smem;
    }
    for (int i = R.lane; i < kp.maxR * R.mstride; i += 64)
        R.mbox[i] = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = R.lane + 64 * k;
        if (p < R.P) {
            const long long o = (long long)r * R.P + p;
            R.nr[k] = kp.nr[o];
            R.rate[k] = kp.rates[o];
            R.integ[k] = kp.integ[o];
            R.pstat[k] = kp.pstat[o];
        } else {
            R.nr[k] = 0; R.rate[k] = 0.0; R.integ[k] = 0.0; R.pstat[k] = 0;
        }
    }
    __syncthreads();
}

template <int PPL, int BITS, bool LDS_LAT>
__device__ __forceinline__ void rep_store(Rep<PPL, BITS> &R, const KParams &kp, int r, unsigned char *smem,
                                          bool lattice_changed)
{
    __syncthreads();
    if (LDS_LAT && lattice_changed) {
        const long long lb = lattice_bytes(R.d.V, BITS);
        uint4 *dst = (uint4 *)(kp.lat_pool + kp.lat_off[r]);
        const uint4 *src = (const uint4 *)smem;
        for (long long i = R.lane; i < lb / 16; i += 64)
            dst[i] = src[i];
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int p = R.lane + 64 * k;
        if (p < R.P) {
            const long long o = (long long)r * R.P + p;
            kp.nr[o] = R.nr[k];
            kp.integ[o] = R.integ[k];
            kp.pstat[o] = R.pstat[k];
        }
    }
}

// do_kmc_steps for one replica per workgroup (proclist_generic_subroutines.mpy:1-44)
template <int PPL, int BITS, bool LDS_LAT>
__global__ __launch_bounds__(64) void kmc_steps_kernel(KParams kp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r = kp.rep_ids[blockIdx.x];
    if (kp.status[r] != ST_OK) {
        if (threadIdx.x == 0)
            kp.executed[blockIdx.x] = 0;
        return;
    }
    Rep<PPL, BITS> R;
    rep_load<PPL, BITS, LDS_LAT>(R, kp, r, smem);
    const int lane = R.lane, P = R.P;
    long long n = kp.nsteps;
    if (kp.budget)
        n = min(n, kp.budget[blockIdx.x]);
    double t = kp.ktime[r], dt_last = kp.kdt[r];
    long long ks = kp.kstep[r];
    unsigned long long ctr = kp.rngctr[r];
    const unsigned seed = kp.seed[r], stream = kp.stream[r];
    int st = ST_OK;
    long long done = 0;
    while (done < n) {
        double c[PPL], acc[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            c[k] = (double)R.nr[k] * R.rate[k];
        const double total = accumulate<PPL>(c, acc, P, lane);
        if (!(total > 0.0)) {
            st = ST_ABSORBING;
            break;
        }
        double u[3];
        step_uniforms(seed, stream, ctr, u);
        ctr += 1;
        // update_clocks (base.mpy:1123-1161)
        const double dt = -log(u[0]) / total;
        t = t + dt;
        dt_last = dt;
        ks += 1;
        // update_integ_rate (base.mpy:626-645): (nr*rate)*dt, pre-step nr
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            R.integ[k] = R.integ[k] + c[k] * dt;
        // determine_procsite (base.mpy:1075-1120)
        const int pick = interval_search<PPL>(acc, u[1] * total, P, lane);
        const int pk = pick >> 6, pl = pick & 63;
        const int nrp = readlane_i(slot_get<PPL>(R.nr, pk), pl);
        int kk = (int)(1.0 + u[2] * (double)nrp);
        kk = kk > nrp ? nrp : kk;
        int sv = 0;
        if (lane == pl) {
            sv = R.sites(pick)[kk];
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (k == pk) R.pstat[k] += 1;
        }
        const int site = readlane_i(sv, pl);
        done += 1;
        st = run_proc<PPL, BITS>(R, pick + 1, site);
        if (st)
            break;
    }
    rep_store<PPL, BITS, LDS_LAT>(R, kp, r, smem, true);
    if (lane == 0) {
        kp.ktime[r] = t;
        kp.kdt[r] = dt_last;
        kp.kstep[r] = ks;
        kp.rngctr[r] = ctr;
        kp.status[r] = st;
        kp.executed[blockIdx.x] = done;
        if (kp.budget)
            kp.budget[blockIdx.x] -= done;
    }
}

// touchups, run_proc_nr, get_next_kmc_step (one replica per workgroup)
template <int PPL, int BITS, bool LDS_LAT>
__global__ __launch_bounds__(64) void kmc_misc_kernel(KParams kp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int r = kp.rep_ids[blockIdx.x];
    Rep<PPL, BITS> R;
    rep_load<PPL, BITS, LDS_LAT>(R, kp, r, smem);
    const int lane = R.lane, P = R.P;
    const int *prog = R.prog;
    int st = kp.status[r];
    bool changed = false;
    if (kp.mode == MODE_INIT || kp.mode == MODE_ADJUST || kp.mode == MODE_TOUCHUP) {
        const bool lat_int = prog[H_BACKEND] == 1;
        const int nsite = lat_int ? 1 : R.d.spuck;
        auto touch = [&](int x, int y, int z) -> int {
            for (int n = 1; n <= nsite; ++n) {
                const int seg = prog[prog[H_TOUCHUP] + n - 1];
                const int e = run_segment<PPL, BITS>(R, seg, x, y, z);
                if (e)
                    return e;
            }
            return 0;
        };
        if (kp.mode == MODE_INIT) {
            for (int z = 0; z < R.d.Z && !st; ++z)
                for (int y = 0; y < R.d.Y && !st; ++y)
                    for (int x = 0; x < R.d.X && !st; ++x)
                        st = touch(x, y, z);
        } else if (kp.mode == MODE_ADJUST) {
            for (int x = 0; x < R.d.X && !st; ++x)
                for (int y = 0; y < R.d.Y && !st; ++y)
                    for (int z = 0; z < R.d.Z && !st; ++z)
                        st = touch(x, y, z);
        } else {
            const int seg = prog[prog[H_TOUCHUP] + (lat_int ? 0 : kp.arg3 - 1)];
            st = run_segment<PPL, BITS>(R, seg, kp.arg0, kp.arg1, kp.arg2);
        }
    } else if (kp.mode == MODE_RUN_PROC) {
        const int proc = kp.arg0, site = kp.arg1;
        const int pk = (proc - 1) >> 6, pl = (proc - 1) & 63;
        if (lane == pl) {
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (k == pk) R.pstat[k] += 1;
        }
        st = run_proc<PPL, BITS>(R, proc, site);
        changed = true;
    } else if (kp.mode == MODE_NEXT_STEP) {
        double c[PPL], acc[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            c[k] = (double)R.nr[k] * R.rate[k];
        const double total = accumulate<PPL>(c, acc, P, lane);
        if (!(total > 0.0)) {
            st = ST_ABSORBING;
        } else {
            double u[3];
            const unsigned long long ctr = kp.rngctr[r];
            step_uniforms(kp.seed[r], kp.stream[r], ctr, u);
            const int pick = interval_search<PPL>(acc, u[1] * total, P, lane);
            const int pk = pick >> 6, pl = pick & 63;
            const int nrp = readlane_i(slot_get<PPL>(R.nr, pk), pl);
            int kk = (int)(1.0 + u[0] * (double)nrp);
            kk = kk > nrp ? nrp : kk;
            int sv = 0;
            if (lane == pl)
                sv = R.sites(pick)[kk];
            const int site = readlane_i(sv, pl);
            if (lane == 0) {
                kp.out[0] = pick + 1;
                kp.out[1] = site;
                kp.rngctr[r] = ctr + 1;
            }
        }
    }
    rep_store<PPL, BITS, LDS_LAT>(R, kp, r, smem, changed);
    if (lane == 0)
        kp.status[r] = st;
}

// default species of every site (initialize_state)
template <int BITS>
__global__ void kmc_fill_default_kernel(KParams kp)
{
    const int r = kp.rep_ids[blockIdx.x];
    const int4 dd = kp.dims[r];
    Lattice<BITS> lat;
    lat.p = kp.lat_pool + kp.lat_off[r];
    const int *sdef = kp.prog + kp.prog[H_SITEDEF];
    const long long lb = lattice_bytes(dd.w, BITS);
    if (BITS == 8) {
        for (long long s = threadIdx.x; s < lb; s += blockDim.x)
            lat.p[s] = (unsigned char)(signed char)(s < dd.w ? sdef[s % kp.spuck] : 0);
    } else {
        const int per = 8 / BITS;
        for (long long b = threadIdx.x; b < lb; b += blockDim.x) {
            unsigned v = 0;
            for (int i = 0; i < per; ++i) {
                const long long s = b * per + i;
                const unsigned code = s < dd.w ? (unsigned)(sdef[s % kp.spuck] + 1) : 0u;
                v |= code << (i * BITS);
            }
            lat.p[b] = (unsigned char)v;
        }
    }
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
};

} // namespace

struct kmcg_model {
    std::vector<int> blob;
    std::vector<int> prog;
    int P = 0, S = 0, spuck = 0, dim = 0, backend = 0, nullsp = -1, maxoff = 0, maxR = 1;
};

struct kmcg_batch {
    const kmcg_model *m = nullptr;
    int device = 0;
    int n = 0;
    int bits = 8;
    int lds_mode = -1;
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
    double *occ = nullptr;
    float last_ms = 0.f;
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
        if (m->nullsp != -1)
            return fail(KMCG_EUNSUPPORTED, "multi-lattice models are not supported");

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
                    if (old < 0 || old >= m->S || nw < 0 || nw >= m->S)
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
                const int nupd = W(base + 2), uoff = W(base + 3);
                for (int i = 0; i < nupd; ++i)
                    for (int q = 0; q < 6; ++q)
                        add.writes.push_back(W(uoff + 6 * i + q));
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

        // flatten
        std::vector<int> &pg = m->prog;
        pg.assign(HDR_WORDS, 0);
        pg[H_P] = m->P; pg[H_S] = m->S; pg[H_SPUCK] = m->spuck; pg[H_DIM] = m->dim;
        pg[H_BACKEND] = m->backend; pg[H_NULL] = m->nullsp; pg[H_MAXOFF] = m->maxoff;
        pg[H_MAXR] = m->maxR;
        pg[H_SITEDEF] = (int)pg.size();
        for (int n = 0; n < m->spuck; ++n)
            pg.push_back(W(W(8) + n));
        pg[H_NSEG] = (int)segs.size();
        pg[H_SEG] = (int)pg.size();
        pg.resize(pg.size() + SEG_W * segs.size(), 0);
        for (size_t si = 0; si < segs.size(); ++si) {
            SegH &sg = segs[si];
            const int sb = pg[H_SEG] + SEG_W * (int)si;
            pg[sb + 0] = (int)sg.writes.size() / WR_W;
            pg[sb + 1] = (int)pg.size();
            pg.insert(pg.end(), sg.writes.begin(), sg.writes.end());
            pg[sb + 2] = (int)sg.ops.size();
            const int ob = (int)pg.size();
            pg[sb + 3] = ob;
            pg[sb + 4] = sg.nrounds;
            pg.resize(pg.size() + OP_W * sg.ops.size(), 0);
            for (size_t j = 0; j < sg.ops.size(); ++j) {
                const OpH &op = sg.ops[j];
                const int o = ob + OP_W * (int)j;
                pg[o + 0] = op.kind; pg[o + 1] = op.round; pg[o + 2] = op.arg;
                pg[o + 3] = op.dx; pg[o + 4] = op.dy; pg[o + 5] = op.dz; pg[o + 6] = op.n;
                pg[o + 7] = (int)op.conds.size() / CD_W;
                pg[o + 8] = (int)pg.size();
                pg.insert(pg.end(), op.conds.begin(), op.conds.end());
            }
        }
        pg[H_PROG] = (int)pg.size();
        pg.resize(pg.size() + 2 * m->P, 0);
        for (int p = 0; p < m->P; ++p) {
            pg[pg[H_PROG] + 2 * p] = (int)proc_items[p].size() / IT_W;
            pg[pg[H_PROG] + 2 * p + 1] = (int)pg.size();
            for (size_t i = 0; i < proc_items[p].size(); i += IT_W)
                if (proc_items[p][i] < 0 || proc_items[p][i] >= (int)segs.size())
                    return fail(KMCG_EINVAL, "process program names a bad segment");
            pg.insert(pg.end(), proc_items[p].begin(), proc_items[p].end());
        }
        pg[H_TOUCHUP] = (int)pg.size();
        pg.insert(pg.end(), touchup.begin(), touchup.end());
        pg[H_NGROUPS] = (int)nli_nodes.size();
        pg[H_NLI] = (int)pg.size();
        pg.resize(pg.size() + nli_nodes.size(), 0);
        for (size_t g = 0; g < nli_nodes.size(); ++g) {
            const int base = (int)pg.size();
            pg[pg[H_NLI] + g] = base;
            std::vector<int> nodes = nli_nodes[g];
            // child offsets become absolute
            size_t pos = 0;
            while (pos < nodes.size()) {
                const int nc = nodes[pos + 4];
                if (nc == -1) {
                    pos += 6;
                    continue;
                }
                for (int k = 0; k < nc; ++k)
                    if (nodes[pos + 6 + 3 * k] == 0)
                        nodes[pos + 7 + 3 * k] += base;
                pos += 5 + 3 * nc;
            }
            pg.insert(pg.end(), nodes.begin(), nodes.end());
        }
        // every nli tree: nodes in bounds, children strictly after their
        // parent (so the device walk terminates), leaves name processes
        for (size_t g = 0; g < nli_nodes.size(); ++g) {
            const int base = pg[pg[H_NLI] + g];
            const int end = base + (int)nli_nodes[g].size();
            int pos = base;
            while (pos < end) {
                if (pos + 5 > end)
                    return fail(KMCG_EINVAL, "truncated nli node");
                const int nc = pg[pos + 4];
                if (nc == -1) {
                    if (pos + 6 > end || pg[pos + 5] < 0 || pg[pos + 5] > m->P)
                        return fail(KMCG_EINVAL, "bad nli leaf");
                    pos += 6;
                    continue;
                }
                if (nc < 0 || pos + 5 + 3 * nc > end || pg[pos + 3] < 1 || pg[pos + 3] > m->spuck)
                    return fail(KMCG_EINVAL, "bad nli node");
                for (int k = 0; k < nc; ++k) {
                    const int kind = pg[pos + 6 + 3 * k], tgt = pg[pos + 7 + 3 * k];
                    if (kind == 1 && (tgt < 0 || tgt > m->P))
                        return fail(KMCG_EINVAL, "bad nli leaf value");
                    if (kind == 0 && (tgt <= pos || tgt >= end))
                        return fail(KMCG_EINVAL, "bad nli child offset");
                    if (kind != 0 && kind != 1)
                        return fail(KMCG_EINVAL, "bad nli case kind");
                }
                pos += 5 + 3 * nc;
            }
        }
        for (size_t si = 0; si < segs.size(); ++si)
            for (size_t w = 0; w < segs[si].writes.size(); w += WR_W) {
                const int n = segs[si].writes[w + 3], nw = segs[si].writes[w + 5];
                if (n < 1 || n > m->spuck || nw < 0 || nw >= m->S)
                    return fail(KMCG_EINVAL, "bad lattice write");
            }
        for (int t : touchup)
            if (t < 0 || t >= (int)segs.size())
                return fail(KMCG_EINVAL, "bad touchup segment");
        pg[H_WORDS] = (int)pg.size();
    } catch (const std::exception &e) {
        return fail(KMCG_EINVAL, std::string("malformed model blob: ") + e.what());
    }
    return KMCG_OK;
}

int ppl_of(int P) { return P <= 64 ? 1 : (P <= 128 ? 2 : 4); }

KParams make_params(kmcg_batch *b)
{
    KParams kp{};
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
    kp.executed = b->executed;
    kp.out = b->out;
    kp.P = b->m->P;
    kp.S = b->m->S;
    kp.spuck = b->m->spuck;
    kp.dim = b->m->dim;
    kp.maxR = b->m->maxR;
    return kp;
}

// LDS bytes a launch over `ids` needs, and whether the lattice goes to LDS.
void launch_lds(kmcg_batch *b, const int *ids, int n, bool &lds_lat, size_t &bytes, int &lat_bytes)
{
    long long maxlb = 0;
    for (int i = 0; i < n; ++i)
        maxlb = std::max(maxlb, lattice_bytes(b->V[ids[i]], b->bits));
    const size_t mbox = (size_t)b->m->maxR * 64 * ppl_of(b->m->P) * sizeof(int);
    const bool fits = (size_t)maxlb + mbox <= (size_t)b->max_lds;
    lds_lat = b->lds_mode == 1 ? fits : (b->lds_mode == 0 ? false : fits);
    lat_bytes = lds_lat ? (int)maxlb : 0;
    bytes = (size_t)lat_bytes + mbox;
}

template <int PPL, int BITS>
hipError_t launch_t(bool steps, bool lds_lat, int grid, size_t lds, hipStream_t st, const KParams &kp)
{
    if (steps) {
        if (lds_lat)
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, true>), dim3(grid), dim3(64), lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_steps_kernel<PPL, BITS, false>), dim3(grid), dim3(64), lds, st, kp);
    } else {
        if (lds_lat)
            hipLaunchKernelGGL((kmc_misc_kernel<PPL, BITS, true>), dim3(grid), dim3(64), lds, st, kp);
        else
            hipLaunchKernelGGL((kmc_misc_kernel<PPL, BITS, false>), dim3(grid), dim3(64), lds, st, kp);
    }
    return hipGetLastError();
}

template <int PPL>
hipError_t launch_b(int bits, bool steps, bool lds_lat, int grid, size_t lds, hipStream_t st, const KParams &kp)
{
    if (bits == 2)
        return launch_t<PPL, 2>(steps, lds_lat, grid, lds, st, kp);
    return launch_t<PPL, 8>(steps, lds_lat, grid, lds, st, kp);
}

hipError_t launch(kmcg_batch *b, bool steps, bool lds_lat, int grid, size_t lds, const KParams &kp)
{
    switch (ppl_of(b->m->P)) {
    case 1: return launch_b<1>(b->bits, steps, lds_lat, grid, lds, b->stream, kp);
    case 2: return launch_b<2>(b->bits, steps, lds_lat, grid, lds, b->stream, kp);
    default: return launch_b<4>(b->bits, steps, lds_lat, grid, lds, b->stream, kp);
    }
}

int check_replica(kmcg_batch *b, int r)
{
    if (!b || r < 0 || r >= b->n)
        return fail(KMCG_EINVAL, "replica index out of range");
    return KMCG_OK;
}

// run the misc kernel for replicas [ids] synchronously
int run_misc(kmcg_batch *b, const std::vector<int> &ids, int mode, int a0, int a1, int a2, int a3)
{
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpyAsync(b->rep_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    bool lds_lat;
    size_t lds;
    int lat_bytes;
    launch_lds(b, ids.data(), (int)ids.size(), lds_lat, lds, lat_bytes);
    KParams kp = make_params(b);
    kp.mode = mode;
    kp.arg0 = a0; kp.arg1 = a1; kp.arg2 = a2; kp.arg3 = a3;
    kp.lat_lds = lat_bytes;
    HIPCHK(launch(b, false, lds_lat, (int)ids.size(), lds, kp));
    HIPCHK(hipStreamSynchronize(b->stream));
    return KMCG_OK;
}

template <class T>
int copy_out(kmcg_batch *b, T *dst, const T *src, size_t count)
{
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
    return KMCG_OK;
}

template <class T>
int copy_in(kmcg_batch *b, T *dst, const T *src, size_t count)
{
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return KMCG_OK;
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
    if (!blob || !out || nwords < HDR_WORDS)
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
            return fail(KMCG_EINVAL, "system size smaller than the model's largest offset ("
                                         + std::to_string(m->maxoff) + ")");
        if ((long long)X * Y * Z * m->spuck >= (1LL << 28))
            return fail(KMCG_EUNSUPPORTED, "more than 2^28 sites per replica");
    }
    return KMCG_OK;
}

int kmcg_batch_footprint(const kmcg_model *m, int32_t n, const int32_t *sizes, int64_t *bytes)
{
    int rc = sizes_ok(m, n, sizes);
    if (rc)
        return rc;
    const int bits = (m->S + 1 <= 4) ? 2 : 8;
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
                    b->rngctr, b->status, b->executed, b->budget, b->rep_ids, b->out, b->occ};
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
    b->bits = (m->S + 1 <= 4) ? 2 : 8;
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
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        b->max_lds = (int)prop.sharedMemPerBlock;
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
    BALLOC(b->executed, n);
    BALLOC(b->budget, n);
    BALLOC(b->rep_ids, n);
    BALLOC(b->out, 16);
    BALLOC(b->occ, (size_t)n * m->S * m->spuck);
#undef BALLOC
    auto H2D = [&](void *d, const void *h, size_t bytes) { return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess; };
    if (!H2D(b->prog, m->prog.data(), m->prog.size() * sizeof(int)) ||
        !H2D(b->dims, dims.data(), n * sizeof(int4)) ||
        !H2D(b->lat_off_d, b->lat_off.data(), n * sizeof(long long)) ||
        !H2D(b->avail_off_d, b->avail_off.data(), n * sizeof(long long)))
        return bail(fail(KMCG_EHIP, "copy to device failed"));
    if (hipMemset(b->rates, 0, n * P * sizeof(double)) != hipSuccess ||
        hipMemset(b->status, 0, n * sizeof(int)) != hipSuccess ||
        hipMemset(b->avail_pool, 0, b->avail_total * sizeof(int)) != hipSuccess ||
        hipMemset(b->lat_pool, 0, b->lat_total) != hipSuccess)
        return bail(fail(KMCG_EHIP, "memset failed"));
    *out = b;
    return KMCG_OK;
}

int kmcg_set_launch(kmcg_batch *b, int32_t lattice_in_lds)
{
    if (!b || lattice_in_lds < -1 || lattice_in_lds > 1)
        return fail(KMCG_EINVAL, "bad argument");
    b->lds_mode = lattice_in_lds;
    return KMCG_OK;
}

int kmcg_init(kmcg_batch *b, int32_t r0, int32_t n, const uint32_t *seeds, const uint32_t *streams)
{
    if (!b || r0 < 0 || n < 1 || r0 + n > b->n || !seeds)
        return fail(KMCG_EINVAL, "bad replica range");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    const long long P = b->m->P;
    std::vector<unsigned> st(n, 0u);
    if (streams)
        st.assign(streams, streams + n);
    HIPCHK(hipMemcpy(b->seed + r0, seeds, n * sizeof(unsigned), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->strm + r0, st.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->nr + r0 * P, 0, n * P * sizeof(int)));
    HIPCHK(hipMemset(b->integ + r0 * P, 0, n * P * sizeof(double)));
    HIPCHK(hipMemset(b->pstat + r0 * P, 0, n * P * sizeof(long long)));
    HIPCHK(hipMemset(b->ktime + r0, 0, n * sizeof(double)));
    HIPCHK(hipMemset(b->kdt + r0, 0, n * sizeof(double)));
    HIPCHK(hipMemset(b->kstep + r0, 0, n * sizeof(long long)));
    HIPCHK(hipMemset(b->rngctr + r0, 0, n * sizeof(unsigned long long)));
    HIPCHK(hipMemset(b->status + r0, 0, n * sizeof(int)));
    const long long a0 = b->avail_off[r0];
    const long long a1 = (r0 + n < b->n) ? b->avail_off[r0 + n] : b->avail_total;
    HIPCHK(hipMemset(b->avail_pool + a0, 0, (a1 - a0) * sizeof(int)));
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i)
        ids[i] = r0 + i;
    HIPCHK(hipMemcpy(b->rep_ids, ids.data(), n * sizeof(int), hipMemcpyHostToDevice));
    KParams kp = make_params(b);
    if (b->bits == 2)
        hipLaunchKernelGGL((kmc_fill_default_kernel<2>), dim3(n), dim3(256), 0, b->stream, kp);
    else
        hipLaunchKernelGGL((kmc_fill_default_kernel<8>), dim3(n), dim3(256), 0, b->stream, kp);
    HIPCHK(hipGetLastError());
    return run_misc(b, ids, MODE_INIT, 0, 0, 0, 0);
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
    long long off = 0;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) {
        const int r = r0 + i;
        const int V = b->V[r];
        for (int s = 0; s < V; ++s)
            if (config[off + s] < 0 || config[off + s] >= b->m->S)
                return fail(KMCG_EINVAL, "configuration holds an unknown species");
        std::vector<unsigned char> packed(lattice_bytes(V, b->bits), 0);
        if (b->bits == 8) {
            for (int s = 0; s < V; ++s)
                packed[s] = (unsigned char)config[off + s];
        } else {
            for (int s = 0; s < V; ++s)
                packed[(s * 2) >> 3] |= (unsigned char)(((config[off + s] + 1) & 3) << ((s * 2) & 7));
        }
        const int rc = copy_in(b, b->lat_pool + b->lat_off[r], packed.data(), packed.size());
        if (rc)
            return rc;
        off += V;
        ids[i] = r;
    }
    return run_misc(b, ids, MODE_ADJUST, 0, 0, 0, 0);
}

int kmcg_run(kmcg_batch *b, const int32_t *replicas, int32_t n, int64_t steps, int64_t *budget)
{
    if (!b || !replicas || n < 1 || n > b->n || steps < 0)
        return fail(KMCG_EINVAL, "bad run arguments");
    for (int i = 0; i < n; ++i)
        if (replicas[i] < 0 || replicas[i] >= b->n)
            return fail(KMCG_EINVAL, "replica index out of range");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpyAsync(b->rep_ids, replicas, n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    if (budget)
        HIPCHK(hipMemcpyAsync(b->budget, budget, n * sizeof(long long), hipMemcpyHostToDevice, b->stream));
    bool lds_lat;
    size_t lds;
    int lat_bytes;
    launch_lds(b, replicas, n, lds_lat, lds, lat_bytes);
    KParams kp = make_params(b);
    kp.nsteps = steps;
    kp.budget = budget ? b->budget : nullptr;
    kp.lat_lds = lat_bytes;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(launch(b, true, lds_lat, n, lds, kp));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    if (budget) {
        HIPCHK(hipMemcpyAsync(budget, b->budget, n * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return KMCG_OK;
}

int kmcg_sync(kmcg_batch *b)
{
    if (!b)
        return fail(KMCG_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess)
        b->last_ms = ms;
    return KMCG_OK;
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
    if (site < 1 || site > b->V[r] || new_sp < 0 || new_sp >= b->m->S)
        return fail(KMCG_EINVAL, "site or species out of range");
    std::vector<int8_t> lat(b->V[r]);
    rc = kmcg_get_lattice(b, r, lat.data());
    if (rc)
        return rc;
    if (lat[site - 1] != old_sp)
        return fail(KMCG_EINVAL, "replace_species: site holds another species");
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

} // extern "C"
