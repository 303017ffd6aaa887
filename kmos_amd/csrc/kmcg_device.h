// kmcg_device.h -- device-side data layout and step machinery of the engine.
//
// One wavefront (64 lanes) runs one replica.  The lattice of the replica
// sits bit-packed in LDS; the available-site lists (kmos's avail_sites) sit
// in HBM as two int32 arrays per process, sites[p][0..V] and back[p][0..V]
// (1-based, slot 0 unused).  Lane (p % 64) owns process p: it holds
// nr_of_sites, the rate, integ_rate and procstat of p in registers and is
// the only lane that ever touches sites[p] / back[p], so the order of the
// bookkeeping operations on one process is plain program order.
//
// Every generated routine of the reference (put/take/touchup for
// local_smart, run_proc_<group>/touchup_cell for lat_int) is a "segment":
// lattice writes followed by an ordered op list.  Ops are evaluated
// lane-parallel (conditions, nli case trees), routed through an LDS mailbox
// to the lane owning their process, and applied in rounds precomputed on the
// host so that ops in one round touch pairwise different processes and ops
// on one process keep the reference order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmcg {

// ---- device program layout (int32 words; built by the host from the blob)
enum : int {
    H_P = 0, H_S, H_SPUCK, H_DIM, H_BACKEND, H_NULL, H_MAXOFF, H_MAXR,
    H_SEG, H_WRITES, H_OPS, H_CONDS, H_PROG, H_ITEMS, H_TOUCHUP, H_NLI,
    H_NGROUPS, H_SITEDEF, H_WORDS, H_NSEG,
    HDR_WORDS = 32
};
// segment: [n_writes, write_off, n_ops, op_off, n_rounds, 0]
constexpr int SEG_W = 6;
// write: [dx, dy, dz, n, old_mask, new_species]
constexpr int WR_W = 6;
// op: [kind, round, arg, dx, dy, dz, n, n_cond, cond_off, 0]
constexpr int OP_W = 10;
// cond: [dx, dy, dz, n, species]
constexpr int CD_W = 5;
// process program: [n_items, item_off]; item: [segment, dx, dy, dz]
constexpr int IT_W = 4;

enum : int { OP_DEL_IF = 1, OP_ADD_IF = 2, OP_DEL_NLI = 3, OP_ADD_NLI = 4 };
// mailbox codes (low 2 bits of an entry; site+1 in the upper bits)
enum : int { MB_DEL_IF = 1, MB_DEL_STRICT = 2, MB_ADD = 3 };

enum : int { ST_OK = 0, ST_ABSORBING = 1, ST_MISMATCH = 2, ST_DEL_ABSENT = 3, ST_CORRUPT = 4 };

struct Dims {
    int X, Y, Z, V, spuck, dim;
};

// Kernel parameters (one struct, passed by value).
struct KParams {
    const int *prog;
    const int *rep_ids;
    const int4 *dims;              // X, Y, Z, V
    const long long *lat_off;      // bytes into lat_pool
    const long long *avail_off;    // words into avail_pool
    const unsigned *seed;
    const unsigned *stream;
    unsigned char *lat_pool;
    int *avail_pool;
    int *nr;
    double *rates;
    double *integ;
    long long *pstat;
    double *ktime;
    double *kdt;
    long long *kstep;
    unsigned long long *rngctr;
    int *status;
    long long *executed;
    long long *budget;             // optional, per launch slot
    long long nsteps;
    int P, S, spuck, dim, maxR;
    int lat_lds;                   // LDS bytes reserved for the lattice (0: lattice in HBM)
    int mode;                      // misc kernel mode
    int arg0, arg1, arg2, arg3;    // misc kernel arguments
    int *out;                      // misc kernel outputs
};

// ---- helpers --------------------------------------------------------------
__device__ __forceinline__ int wrap1(int a, int n)
{
    a += (a < 0) ? n : 0;
    a -= (a >= n) ? n : 0;
    return a;
}

// 0-based site index of cell (x, y, z) + offset, site type n (1-based);
// periodic wrap as lattice2nr (kmos/fortran_src/lattice.mpy:146-173).
__device__ __forceinline__ int site0(const Dims &d, int x, int y, int z,
                                     int dx, int dy, int dz, int n)
{
    int xx = wrap1(x + dx, d.X);
    int yy = d.dim >= 2 ? wrap1(y + dy, d.Y) : 0;
    int zz = d.dim >= 3 ? wrap1(z + dz, d.Z) : 0;
    return d.spuck * (xx + d.X * (yy + d.Y * zz)) + n - 1;
}

__device__ __forceinline__ int readlane_i(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    long long b = __double_as_longlong(v);
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ void philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3,
                                       unsigned k0, unsigned k1, unsigned out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        unsigned n0 = hi1 ^ c1 ^ k0;
        unsigned n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double u52(unsigned lo, unsigned hi)
{
    unsigned long long x = ((unsigned long long)hi << 32) | lo;
    return ((double)(x >> 12) + 0.5) * 0x1p-52;
}

// The three uniforms of one step: time, proc, site.
__device__ __forceinline__ void step_uniforms(unsigned seed, unsigned stream,
                                              unsigned long long step, double u[3])
{
    unsigned o[4];
    philox((unsigned)step, (unsigned)(step >> 32), 0u, 0u, seed, stream, o);
    u[0] = u52(o[0], o[1]);
    u[1] = u52(o[2], o[3]);
    philox((unsigned)step, (unsigned)(step >> 32), 1u, 0u, seed, stream, o);
    u[2] = u52(o[0], o[1]);
}

// Bit-packed lattice: BITS = 8 stores the species as int8; BITS = 2 or 4
// store species+1 (0 = null species -1).
template <int BITS>
struct Lattice {
    unsigned char *p;
    __device__ __forceinline__ int get(int s) const
    {
        if (BITS == 8)
            return (int)(signed char)p[s];
        int bit = s * BITS;
        return (int)((p[bit >> 3] >> (bit & 7)) & ((1 << BITS) - 1)) - 1;
    }
    __device__ __forceinline__ void set(int s, int v)
    {
        if (BITS == 8) {
            p[s] = (unsigned char)(signed char)v;
            return;
        }
        int bit = s * BITS;
        unsigned char m = (unsigned char)(((1 << BITS) - 1) << (bit & 7));
        unsigned char b = p[bit >> 3];
        p[bit >> 3] = (unsigned char)((b & ~m) | ((((unsigned)(v + 1)) << (bit & 7)) & m));
    }
};

__host__ __device__ inline long long lattice_bytes(int V, int bits)
{
    long long b = ((long long)V * bits + 7) / 8;
    return (b + 15) & ~15LL;
}

} // namespace kmcg
