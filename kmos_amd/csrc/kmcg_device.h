// kmcg_device.h -- device-side data layout of the kMC engine.
//
// One wavefront (64 lanes) runs one replica; a workgroup holds up to four
// replica-waves that share one copy of the model's op program in LDS.  The
// lattice of each replica sits bit-packed in LDS; the available-site lists
// (kmos's avail_sites) sit in HBM as two int32 arrays per process,
// sites[p][0..V] and back[p][0..V] (1-based, slot 0 unused).  Lane (p % 64)
// owns process p: it keeps nr_of_sites, the rate, integ_rate and procstat of
// p in registers and is the only lane that ever reads or writes sites[p] /
// back[p], so the order of the bookkeeping operations on one process is the
// lane's program order.
//
// Every generated routine of the reference (put/take/touchup for
// local_smart, run_proc_<group>/touchup_cell for lat_int) is a "segment":
// lattice writes followed by an ordered op list.  Ops are evaluated
// lane-parallel (conditions, nli case trees) and posted through LDS mailbox
// rows to the lane owning their process, in rounds precomputed on the host
// so that one process's ops keep the reference order.  The segments of one
// executed process post to consecutive rows; the owner lane then replays
// all its entries of the step from one batch of HBM reads per block of
// rows (replay_block / apply_rows in kmcg_engine.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmcg {

// ---- packed op program (int32 words, copied to LDS) -----------------------
// segment (6 words): n_writes, write_off, n_ops | n_rounds << 16 | as_add << 24,
//                    op_off, cond_base, 0
//                    (as_add: the ops are another segment's nli deletions,
//                    applied here as additions)
// write   (2 words): pack3(dx,dy,dz) | n << 24,  old_mask | new_species << 16
// op      (2 words): kind | round << 3 | n << 11 | n_cond << 17 | arg << 23,
//                    dx & 127 | (dy & 127) << 7 | (dz & 127) << 14 | cond_rel << 21
// cond    (1 word) : (dx+32) | (dy+32) << 6 | (dz+32) << 12 | n << 18 | species << 24
// process (2 words): n_items, item_off;  item (2 words): segment, pack3(dx,dy,dz)
// touchup (1 word per site type, 1 for lat_int): segment
// nli     (1 word per group): node offset; node: pack3 | n << 24, n_cases,
//                    (species_mask, kind, target) x n_cases  (kind 0 child, 1 leaf)
//         leaf node: 0, -1, value
// nli table (1 word per group, -1 = none): K | S << 8, K coordinates
//         (pack3 | n << 24), then S^K bytes: the variant for the species
//         index sum_k species_k * S^k
constexpr int SEG_W = 6, WR_W = 2, OP_W = 2, IT_W = 2;

__device__ __forceinline__ int sext7(int v) { return (v ^ 0x40) - 0x40; }

enum : int { OP_DEL_IF = 1, OP_ADD_IF = 2, OP_DEL_NLI = 3, OP_ADD_NLI = 4 };
// mailbox codes (low 2 bits of an entry; site+1 in the upper bits)
enum : int { MB_DEL_IF = 1, MB_DEL_STRICT = 2, MB_ADD = 3 };

enum : int { ST_OK = 0, ST_ABSORBING = 1, ST_MISMATCH = 2, ST_DEL_ABSENT = 3, ST_CORRUPT = 4 };

constexpr int MAX_WPG = 4;    // replica-waves per workgroup

__host__ __device__ inline int pack3(int dx, int dy, int dz)
{
    return (dx & 0xFF) | ((dy & 0xFF) << 8) | ((dz & 0xFF) << 16);
}
__device__ __forceinline__ int unpack_x(int w) { return (int)(signed char)(w & 0xFF); }
__device__ __forceinline__ int unpack_y(int w) { return (int)(signed char)((w >> 8) & 0xFF); }
__device__ __forceinline__ int unpack_z(int w) { return (int)(signed char)((w >> 16) & 0xFF); }

struct Dims {
    int X, Y, Z, V, spuck, dim;
    int sX, sXY;        // spuck * X, spuck * X * Y (site strides of y and z)
    int moff;           // largest offset component any segment uses
    __device__ __forceinline__ void strides()
    {
        sX = spuck * X;
        sXY = sX * Y;
    }
    // base cells at least moff away from every face: no offset wraps
    __device__ __forceinline__ bool interior(int x, int y, int z) const
    {
        return x >= moff && x < X - moff && (dim < 2 || (y >= moff && y < Y - moff)) &&
               (dim < 3 || (z >= moff && z < Z - moff));
    }
    __device__ __forceinline__ int lin(int x, int y, int z) const
    {
        return __umul24(x, spuck) + __umul24(y, sX) + __umul24(z, sXY);
    }
    // site offset of (dx, dy, dz, n) from a cell, without wrapping
    __device__ __forceinline__ int off(int dx, int dy, int dz, int n) const
    {
        return __mul24(dx, spuck) + __mul24(dy, sX) + __mul24(dz, sXY) + n - 1;
    }
};

// Kernel parameters (one struct, passed by value).
struct KParams {
    const int *prog;               // packed op program in HBM (copied to LDS)
    const int *rep_ids;            // replicas of this launch
    const int4 *dims;              // X, Y, Z, V per replica
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
    const int *lists_ok;           // per replica: lists match the lattice (or null)
    long long *executed;
    long long *budget;             // optional, per launch slot
    int *out;                      // misc kernel outputs
    long long nsteps;
    int nrep;                      // replicas in this launch
    int wpg;                       // replica-waves per workgroup
    int prog_words, prog_lds;      // program size (words) and its LDS bytes
    int wave_lds;                  // LDS bytes per wave (lattice + mailbox)
    int lat_lds;                   // LDS bytes of the lattice part (0: lattice in HBM)
    int P, S, spuck, dim, maxR, backend;
    int moff;                      // largest offset component of the program
    int mrows;                     // mailbox rows per wave
    int merged;                    // bit 0: the segments of a step share one apply;
                                   // bit 1: a lat_int group's segments post in one pass
    int off_seg, off_proc, off_touch, off_nli, off_nlitab;
    double *srec;                  // in-loop sample records (or null)
    int sample_every, srec_cap, srec_width;
    int mode;                      // misc kernel mode
    int arg0, arg1, arg2, arg3;    // misc kernel arguments
    // displacement tracking (msd_hop null: off)
    const int *msd_hop;            // [P][MSD_HOP_W]: hop?, from dx dy dz n, to dx dy dz n
    const double *msd_d;           // [P][3] displacement of each hop
    int *msd_id;                   // id_arr, [V+1] per replica at msd_off
    int *msd_site;                 // site_arr, [V+1] per replica at msd_off
    double *msd_disp;              // displacement, [V+1][3] per replica at 3 * msd_off
    const long long *msd_off;      // per replica
    int *msd_ev;                   // trajectory events (particle, site), [msd_evcap] per replica
    long long *msd_evn;            // events recorded per replica
    long long msd_evcap;           // 0: no trajectory
    // autocorrelation bins (acf_pa null: off), over the tracking arrays above
    const int *acf_pa;             // property_acf, [V+1] per replica at msd_off
    int *acf_po;                   // property_o, [V+1] per replica at msd_off
    double *acf_buf;               // buffer_acf, [V+1] per replica at msd_off
    const double *acf_pp;          // product_property [acf_nt1][acf_nt1]
    int acf_nt1;                   // number of property types + 1
    double *acf_bin;               // config_bin per replica at acf_boff
    int *acf_cnt;                  // counter_write_in_bin per replica at acf_boff
    const long long *acf_boff;
    double *acf_d;                 // [8] per replica: sum(buffer_acf), t_o, wrap_count, t_bin
    int *acf_i;                    // [8] per replica: bin_index, nr_of_bins, extended bins,
                                   // nr_of_ions_before_kmc_step_acf, nr_of_ions
    // phase profile (library built with KMCG_PHASES, else null): per
    // replica 8 counters, shader-clock ticks spent in select, site read,
    // post, replay and the rest of the step, then the steps counted
    unsigned long long *phase;
};

constexpr int PHASE_W = 8;

constexpr int MSD_HOP_W = 9;

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
    const int xx = wrap1(x + dx, d.X);
    const int yy = d.dim >= 2 ? wrap1(y + dy, d.Y) : 0;
    const int zz = d.dim >= 3 ? wrap1(z + dz, d.Z) : 0;
    return d.lin(xx, yy, zz) + n - 1;
}

// site0 from a base cell: IN (base cell interior) skips the wrap, lb is
// the base cell's first site d.lin(x, y, z).  A template parameter, so that
// the caller branches once instead of computing both forms.
template <bool IN>
__device__ __forceinline__ int site_from(const Dims &d, int lb, int x, int y, int z, int dx, int dy, int dz, int n)
{
    if (IN)
        return lb + d.off(dx, dy, dz, n);
    return site0(d, x, y, z, dx, dy, dz, n);
}

// a / d for 0 <= a < 2^31 with m = ceil(2^32 / d) (d > 1): one multiply-high
// and one correction.
struct Magic {
    unsigned m;
    int d;
};
__device__ __forceinline__ Magic make_magic(int d)
{
    Magic g;
    g.d = d;
    g.m = d > 1 ? (0xFFFFFFFFu / (unsigned)d) + 1u : 0u;
    return g;
}
__device__ __forceinline__ int divm(int a, const Magic &g)
{
    if (g.d == 1)
        return a;
    unsigned q = __umulhi((unsigned)a, g.m);
    if (q * (unsigned)g.d > (unsigned)a)
        q -= 1;
    return (int)q;
}

__device__ __forceinline__ int readlane_i(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// LDS ordering among the lanes of one wave: LDS instructions of a wave
// execute in issue order, so only compiler reordering has to be stopped.
// Unlike __syncthreads() this does not wait for outstanding HBM stores.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ void philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3,
                                       unsigned k0, unsigned k1, unsigned out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const unsigned n0 = hi1 ^ c1 ^ k0;
        const unsigned n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double u52(unsigned lo, unsigned hi)
{
    const unsigned long long x = ((unsigned long long)hi << 32) | lo;
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

// Bit-packed lattice: BITS = 8 stores the species as int8; BITS = 2 stores
// species+1 (0 = null species -1).
template <int BITS>
struct Lattice {
    unsigned char *p;
    __device__ __forceinline__ int get(int s) const
    {
        if (BITS == 8)
            return (int)(signed char)p[s];
        const int bit = s * BITS;
        return (int)((p[bit >> 3] >> (bit & 7)) & ((1 << BITS) - 1)) - 1;
    }
    __device__ __forceinline__ void set(int s, int v)
    {
        if (BITS == 8) {
            p[s] = (unsigned char)(signed char)v;
            return;
        }
        const int bit = s * BITS;
        const unsigned char m = (unsigned char)(((1 << BITS) - 1) << (bit & 7));
        const unsigned char b = p[bit >> 3];
        p[bit >> 3] = (unsigned char)((b & ~m) | ((((unsigned)(v + 1)) << (bit & 7)) & m));
    }
    // old -> v at site s; safe against other lanes writing other sites of
    // the same word at the same time (the buffer is 4-byte aligned)
    __device__ __forceinline__ void xset(int s, int old, int v)
    {
        if (BITS == 8) {
            p[s] = (unsigned char)(signed char)v;
            return;
        }
        const int bit = s * BITS;
        const unsigned x = (((unsigned)(old + 1)) ^ ((unsigned)(v + 1))) << (bit & 31);
        atomicXor((unsigned *)p + (bit >> 5), x);
    }
};

__host__ __device__ inline long long lattice_bytes(int V, int bits)
{
    const long long b = ((long long)V * bits + 7) / 8;
    return (b + 15) & ~15LL;
}

} // namespace kmcg
