/*
 * kmc_oracle.c -- CPU restatement of kmos's generated kMC step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * engine in kmos_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * It interprets the same compiled op programs the engine runs (the int32
 * model blob written by kmos_amd/compiler.py) with the bookkeeping rules of
 * the reference Fortran, one replica at a time, strictly sequentially:
 *
 *   del_proc / add_proc / can_do      kmos/fortran_src/base.mpy:211-321
 *   lat_int proc==0 no-op             kmos/fortran_src/base_lat_int.mpy:197-264
 *   update_accum_rate                 base.mpy:603-623
 *   update_integ_rate                 base.mpy:626-645
 *   determine_procsite                base.mpy:1075-1120
 *   update_clocks                     base.mpy:1123-1161
 *   replace_species                   base.mpy:1187-1231
 *   interval_search_real              base.mpy:1234-1338
 *   lattice2nr / nr2lattice           kmos/fortran_src/lattice.mpy:146-210
 *   do_kmc_step(s), get_next_kmc_step proclist_generic_subroutines.mpy:1-110
 *   get_occupation                    proclist_generic_subroutines.mpy:113-158
 *   reload_system (state overwrite)   base.mpy:363-514
 *   displacement tracking (MSD)       base_acf.mpy:217-361,825-891,
 *                                     proclist_generic_subroutines_acf.mpy:143-219
 *   autocorrelation bins (ACF)        base_acf.mpy:382-552,1053-1201,
 *                                     proclist_generic_subroutines_acf.mpy:1-98
 *   initialize_state                  proclist_generic_subroutines.mpy:236-304
 *   put/take/touchup routines         kmos/io/__init__.py:2219-2443 (as data)
 *   lat_int run_proc/nli/touchup_cell kmos/io/__init__.py:1732-2057 (as data)
 *
 * The uniform stream is the counter-based Philox4x32-10 stream shared with
 * the engine and with the injected reference build (oracle/ref_rng.c).
 *
 * Where the reference has undefined behaviour (total rate zero: division by
 * zero and a read past accum_rates; lat_int del_proc of an unregistered
 * process) the oracle stops the replica with a status code instead.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ABSORBING 1
#define ORC_SPECIES_MISMATCH 2
#define ORC_DEL_ABSENT 3
#define ORC_BAD_ARG 4

#define NO_WRITE (-1000)

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ---------------- */
static void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    philox4x32_10(ctr, key, out);
}

/* (x >> 12) + 0.5 scaled by 2^-52: in (0, 1), exact in binary64. */
static double u52(uint32_t lo, uint32_t hi)
{
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return ((double)(x >> 12) + 0.5) * 0x1p-52;
}

/* The three uniforms of step `step`: time, proc, site. */
void orc_step_uniforms(uint32_t seed, uint32_t stream, uint64_t step, double u[3])
{
    uint32_t key[2] = {seed, stream};
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), 0u, 0u};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    u[0] = u52(o[0], o[1]);
    u[1] = u52(o[2], o[3]);
    ctr[2] = 1u;
    philox4x32_10(ctr, key, o);
    u[2] = u52(o[0], o[1]);
}

/* ---------------- state ---------------- */
typedef struct {
    int32_t *blob;
    int64_t nwords;
    int backend, P, S, nullsp, spuck, dim;
    int X, Y, Z, V;
    int8_t *lat;          /* [V+1], 1-based */
    int32_t *sites;       /* [P][V+1]: avail_sites(:,:,1) */
    int32_t *back;        /* [P][V+1]: avail_sites(:,:,2) */
    int32_t *nr;          /* [P] nr_of_sites */
    double *rates, *accum, *integ;
    int64_t *procstat;
    double kmc_time, kmc_time_step;
    int64_t kmc_step;
    uint32_t seed, stream;
    uint64_t rng_ctr;     /* steps' worth of uniforms drawn */
    int status;
    uint64_t touched;     /* avail entries read or written */
    /* displacement tracking (base_acf.mpy:1-30 module variables) */
    int32_t *msd_hops;    /* [P][9] hop table */
    double *msd_d;        /* [P][3] */
    int32_t *msd_id, *msd_site;  /* id_arr(0:V), site_arr(0:V) */
    double *msd_disp;     /* displacement(0:V, 3) */
    int msd_nions;
    int32_t *msd_ev;      /* trajectory events (particle, site) */
    int64_t msd_evcap, msd_evn;
    /* autocorrelation bins (base_acf.mpy module variables) */
    int32_t *acf_pa, *acf_po;    /* property_acf(0:V), property_o(0:V) */
    double *acf_buf, *acf_pp;    /* buffer_acf(0:V), product_property(0:T,0:T) */
    int acf_nt1, acf_nbins, acf_ext, acf_bin, acf_nr_before;
    double *acf_cb;              /* config_bin(1:ext) */
    int32_t *acf_cc;             /* counter_write_in_bin(1:ext) */
    double acf_tbin, acf_to, acf_wrap;
} orc_t;

#define HDR(o, i) ((o)->blob[(i)])
#define SITES(o, p, k) ((o)->sites[(int64_t)(p) * ((o)->V + 1) + (k)])
#define BACK(o, p, s) ((o)->back[(int64_t)(p) * ((o)->V + 1) + (s)])

/* Count of avail_sites entries read or written (the algorithmic bytes of a
 * step are 4 B times this).  Built with -DORC_NO_TOUCH for timing runs, so
 * that the counter costs nothing there. */
#ifdef ORC_NO_TOUCH
#define TOUCH(o, d) ((void)0)
#else
#define TOUCH(o, d) ((o)->touched += (d))
#endif

static int imod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

/* lattice.mpy:146-173 calculate_lattice2nr, with the periodic modulo. */
static int lattice2nr(const orc_t *o, int x, int y, int z, int n)
{
    int xx = imod(x, o->X);
    int yy = o->dim >= 2 ? imod(y, o->Y) : 0;
    int zz = o->dim >= 3 ? imod(z, o->Z) : 0;
    return o->spuck * (xx + o->X * yy + o->X * o->Y * zz) + n;
}

/* lattice.mpy:177-210 calculate_nr2lattice */
static void nr2lattice(const orc_t *o, int nr, int c[4])
{
    int per_plane = o->X * o->Y * o->spuck;
    c[2] = (nr - 1) / per_plane;
    c[1] = (nr - 1 - per_plane * c[2]) / (o->X * o->spuck);
    c[0] = (nr - 1 - o->spuck * (o->X * o->Y * c[2] + o->X * c[1])) / o->spuck;
    c[3] = nr - o->spuck * (o->X * o->Y * c[2] + o->X * c[1] + c[0]);
}

void orc_nr2lattice(const orc_t *o, int nr, int *c) { nr2lattice(o, nr, c); }
int orc_lattice2nr(const orc_t *o, int x, int y, int z, int n) { return lattice2nr(o, x, y, z, n); }

/* base.mpy:211-265 del_proc (proc is 1-based) */
static void del_proc(orc_t *o, int proc, int site)
{
    int p = proc - 1;
    int m = BACK(o, p, site);
    int last = o->nr[p];
    TOUCH(o, 1);                          /* back(site) read */
    if (m < last) {
        int moved = SITES(o, p, last);
        SITES(o, p, m) = moved;
        SITES(o, p, last) = 0;
        BACK(o, p, moved) = m;
        TOUCH(o, 4);
    } else {
        SITES(o, p, m) = 0;
        TOUCH(o, 1);
    }
    BACK(o, p, site) = 0;
    TOUCH(o, 1);
    o->nr[p] -= 1;
}

/* base.mpy:268-301 add_proc */
static void add_proc(orc_t *o, int proc, int site)
{
    int p = proc - 1;
    o->nr[p] += 1;
    SITES(o, p, o->nr[p]) = site;
    BACK(o, p, site) = o->nr[p];
    TOUCH(o, 2);
}

static int can_do(orc_t *o, int proc, int site)
{
    TOUCH(o, 1);
    return BACK(o, proc - 1, site) != 0;
}

/* base.mpy:1187-1231 replace_species; a mismatch stops the replica. */
static int replace_species(orc_t *o, int site, int old, int new_)
{
    if (o->lat[site] != old) {
        o->status = ORC_SPECIES_MISMATCH;
        return -1;
    }
    o->lat[site] = (int8_t)new_;
    return 0;
}

/* ---------------- local_smart routines ---------------- */
static int routine_run(orc_t *o, int r, const int c[4])
{
    const int32_t *rt = o->blob + HDR(o, 11) + 7 * r;
    int site = lattice2nr(o, c[0], c[1], c[2], c[3]);
    if (rt[1] != NO_WRITE) {
        if (replace_species(o, site, rt[1], rt[2]))
            return -1;
    }
    const int32_t *dis = o->blob + rt[4];
    for (int i = 0; i < rt[3]; ++i) {
        const int32_t *d = dis + 5 * i;
        int s = lattice2nr(o, c[0] + d[1], c[1] + d[2], c[2] + d[3], d[4]);
        if (can_do(o, d[0], s)) {
            TOUCH(o, -1);                  /* del re-uses the check's read */
            del_proc(o, d[0], s);
        }
    }
    const int32_t *en = o->blob + rt[6];
    for (int i = 0; i < rt[5]; ++i) {
        const int32_t *e = en + 7 * i;
        const int32_t *cond = o->blob + e[6];
        int ok = 1;
        for (int k = 0; k < e[5] && ok; ++k) {
            const int32_t *q = cond + 5 * k;
            int s = lattice2nr(o, c[0] + q[0], c[1] + q[1], c[2] + q[2], q[3]);
            ok = (o->lat[s] == q[4]);
        }
        if (ok) {
            int s = lattice2nr(o, c[0] + e[1], c[1] + e[2], c[2] + e[3], e[4]);
            add_proc(o, e[0], s);
        }
    }
    return 0;
}

/* ---------------- lat_int ---------------- */
/* nli_<group>(cell): walk the flattened case tree (io:1985-2057). */
static int nli_eval(const orc_t *o, int g, const int cell[3])
{
    const int32_t *gt = o->blob + HDR(o, 14) + 8 * g;
    const int32_t *base = o->blob + gt[4];
    int pos = 0;
    for (;;) {
        const int32_t *nd = base + pos;
        if (nd[4] == -1)
            return nd[5];
        int s = lattice2nr(o, cell[0] + nd[0], cell[1] + nd[1], cell[2] + nd[2], nd[3]);
        int sp = o->lat[s];
        int next = -1;
        for (int k = 0; k < nd[4]; ++k) {
            const int32_t *e = nd + 5 + 3 * k;
            if (sp >= 0 && (e[0] >> sp) & 1) {
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

static int lat_int_del(orc_t *o, int proc, int site)
{
    if (proc <= 0)
        return 0;
    if (BACK(o, proc - 1, site) == 0) {
        o->status = ORC_DEL_ABSENT;
        return -1;
    }
    del_proc(o, proc, site);
    return 0;
}

static void lat_int_add(orc_t *o, int proc, int site)
{
    if (proc > 0)
        add_proc(o, proc, site);
}

/* run_proc_<group>(cell) (io:1793-1983) */
static int lat_int_run_group(orc_t *o, int g, const int cell[3])
{
    const int32_t *gt = o->blob + HDR(o, 14) + 8 * g;
    const int32_t *mod = o->blob + gt[1];
    for (int i = 0; i < gt[0]; ++i) {
        const int32_t *m = mod + 4 * i;
        int cc[3] = {cell[0] + m[1], cell[1] + m[2], cell[2] + m[3]};
        int p = nli_eval(o, m[0], cc);
        if (lat_int_del(o, p, lattice2nr(o, cc[0], cc[1], cc[2], 1)))
            return -1;
    }
    const int32_t *upd = o->blob + gt[3];
    for (int i = 0; i < gt[2]; ++i) {
        const int32_t *u = upd + 6 * i;
        int s = lattice2nr(o, cell[0] + u[0], cell[1] + u[1], cell[2] + u[2], u[3]);
        int cur = o->lat[s];
        if (cur < 0 || !((u[4] >> cur) & 1)) {
            o->status = ORC_SPECIES_MISMATCH;
            return -1;
        }
        o->lat[s] = (int8_t)u[5];
    }
    for (int i = 0; i < gt[0]; ++i) {
        const int32_t *m = mod + 4 * i;
        int cc[3] = {cell[0] + m[1], cell[1] + m[2], cell[2] + m[3]};
        int p = nli_eval(o, m[0], cc);
        lat_int_add(o, p, lattice2nr(o, cc[0], cc[1], cc[2], 1));
    }
    return 0;
}

/* touchup_cell (io:1766-1791) */
static void lat_int_touchup_cell(orc_t *o, const int cell[3])
{
    int site = lattice2nr(o, cell[0], cell[1], cell[2], 1);
    for (int p = 1; p <= o->P; ++p)
        if (can_do(o, p, site)) {
            TOUCH(o, -1);
            del_proc(o, p, site);
        }
    int G = HDR(o, 13);
    for (int g = 0; g < G; ++g)
        lat_int_add(o, nli_eval(o, g, cell), site);
}

/* ---------------- public API ---------------- */
orc_t *orc_create(const int32_t *blob, int64_t nwords, int X, int Y, int Z)
{
    if (nwords < 32 || blob[0] != 0x4B4D4347 || blob[16] != nwords)
        return NULL;
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    o->blob = (int32_t *)malloc(sizeof(int32_t) * nwords);
    memcpy(o->blob, blob, sizeof(int32_t) * nwords);
    o->nwords = nwords;
    o->backend = blob[2];
    o->P = blob[3];
    o->S = blob[4];
    o->nullsp = blob[5];
    o->spuck = blob[6];
    o->dim = blob[7];
    o->X = X;
    o->Y = o->dim >= 2 ? Y : 1;
    o->Z = o->dim >= 3 ? Z : 1;
    o->V = o->X * o->Y * o->Z * o->spuck;
    /* header word 15: smallest lattice extent the model runs on exactly */
    if (o->X < blob[15] || (o->dim >= 2 && o->Y < blob[15]) || (o->dim >= 3 && o->Z < blob[15])) {
        free(o->blob);
        free(o);
        return NULL;
    }
    int64_t pv = (int64_t)o->P * (o->V + 1);
    o->lat = (int8_t *)malloc(o->V + 1);
    o->sites = (int32_t *)calloc(pv, sizeof(int32_t));
    o->back = (int32_t *)calloc(pv, sizeof(int32_t));
    o->nr = (int32_t *)calloc(o->P, sizeof(int32_t));
    o->rates = (double *)calloc(o->P, sizeof(double));
    o->accum = (double *)calloc(o->P, sizeof(double));
    o->integ = (double *)calloc(o->P, sizeof(double));
    o->procstat = (int64_t *)calloc(o->P, sizeof(int64_t));
    for (int s = 0; s <= o->V; ++s)
        o->lat[s] = (int8_t)o->nullsp;
    return o;
}

void orc_free(orc_t *o)
{
    if (!o) return;
    free(o->blob); free(o->lat); free(o->sites); free(o->back); free(o->nr);
    free(o->rates); free(o->accum); free(o->integ); free(o->procstat);
    free(o->msd_hops); free(o->msd_d); free(o->msd_id); free(o->msd_site); free(o->msd_disp);
    free(o->msd_ev);
    free(o->acf_pa); free(o->acf_po); free(o->acf_buf); free(o->acf_pp); free(o->acf_cb); free(o->acf_cc);
    free(o);
}

/* touchup_<layer>_<site>(site) for local_smart, touchup_cell for lat_int */
int orc_touchup(orc_t *o, int x, int y, int z, int n)
{
    if (o->backend == 0) {
        int r = o->blob[HDR(o, 12) + n - 1];
        int c[4] = {x, y, z, n};
        return routine_run(o, r, c);
    }
    int cell[3] = {x, y, z};
    lat_int_touchup_cell(o, cell);
    return 0;
}

/* initialize_state (proclist_generic_subroutines.mpy:236-304); with a
 * configuration, the lattice starts from it instead of the defaults. */
int orc_init_config(orc_t *o, uint32_t seed, uint32_t stream, const int8_t *config)
{
    const int32_t *sdef = o->blob + HDR(o, 8);
    o->seed = seed;
    o->stream = stream;
    o->rng_ctr = 0;
    o->status = ORC_OK;
    o->kmc_time = 0.0;
    o->kmc_time_step = 0.0;
    o->kmc_step = 0;
    memset(o->nr, 0, sizeof(int32_t) * o->P);
    memset(o->integ, 0, sizeof(double) * o->P);
    memset(o->accum, 0, sizeof(double) * o->P);
    memset(o->procstat, 0, sizeof(int64_t) * o->P);
    memset(o->sites, 0, sizeof(int32_t) * (int64_t)o->P * (o->V + 1));
    memset(o->back, 0, sizeof(int32_t) * (int64_t)o->P * (o->V + 1));
    for (int s = 1; s <= o->V; ++s)
        o->lat[s] = config ? config[s - 1] : (int8_t)sdef[(s - 1) % o->spuck];
    /* local_smart touches up the sites of the initial layer only (header
     * word 17: one flag per site type; every site of a one-layer model) */
    const int32_t *tinit = o->blob + HDR(o, 17);
    for (int k = 0; k < o->Z; ++k)
        for (int j = 0; j < o->Y; ++j)
            for (int i = 0; i < o->X; ++i) {
                if (o->backend == 0) {
                    for (int n = 1; n <= o->spuck; ++n)
                        if (tinit[n - 1] && orc_touchup(o, i, j, k, n)) return o->status;
                } else {
                    orc_touchup(o, i, j, k, 0);
                }
            }
    return o->status;
}

int orc_init(orc_t *o, uint32_t seed, uint32_t stream)
{
    return orc_init_config(o, seed, stream, NULL);
}

/* The lattice of KMC_Model._set_configuration followed by
 * _adjust_database (kmos/run/__init__.py:1408-1457): species are replaced
 * site by site, then every site is touched up with x outermost. */
int orc_set_config(orc_t *o, const int8_t *config /* [V] in site order */)
{
    for (int s = 1; s <= o->V; ++s)
        o->lat[s] = config[s - 1];
    for (int x = 0; x < o->X; ++x)
        for (int y = 0; y < o->Y; ++y)
            for (int z = 0; z < o->Z; ++z) {
                if (o->backend == 0) {
                    for (int n = 1; n <= o->spuck; ++n)
                        if (orc_touchup(o, x, y, z, n)) return o->status;
                } else {
                    orc_touchup(o, x, y, z, 0);
                }
            }
    return o->status;
}

void orc_set_rates(orc_t *o, const double *rates)
{
    memcpy(o->rates, rates, sizeof(double) * o->P);
}

/* base.mpy:603-623 */
static void update_accum_rate(orc_t *o)
{
    o->accum[0] = (double)o->nr[0] * o->rates[0];
    for (int i = 1; i < o->P; ++i)
        o->accum[i] = o->accum[i - 1] + (double)o->nr[i] * o->rates[i];
}

/* base.mpy:626-645 */
static void update_integ_rate(orc_t *o)
{
    for (int i = 0; i < o->P; ++i)
        o->integ[i] = o->integ[i] + (double)o->nr[i] * o->rates[i] * o->kmc_time_step;
}

/* base.mpy:1234-1338 interval_search_real, 1-based result.  The caller
 * guarantees arr[n-1] > 0, which keeps the zero-search in bounds. */
static int interval_search_real(const double *arr, int n, double value)
{
    int left = 1, right = n, mid;
    for (;;) {
        mid = (right + left) >> 1;
        if (left >= right) break;
        if (value < arr[mid - 1]) right = mid;
        else left = mid + 1;
    }
    if (arr[mid - 1] == 0.0) {
        for (;;) {
            if (arr[mid - 1] > 0.0) break;
            mid += 1;
        }
    }
    while (mid != 1) {
        if (arr[mid - 2] >= arr[mid - 1]) mid -= 1;
        else break;
    }
    return mid;
}

/* base.mpy:1075-1120 */
static void determine_procsite(orc_t *o, double ran_proc, double ran_site, int *proc, int *site)
{
    *proc = interval_search_real(o->accum, o->P, ran_proc * o->accum[o->P - 1]);
    int n = o->nr[*proc - 1];
    int k = (int)(1.0 + ran_site * (double)n);
    if (k > n) k = n;
    *site = SITES(o, *proc - 1, k);
    TOUCH(o, 1);
}

/* run_proc_nr for both backends */
int orc_run_proc_nr(orc_t *o, int proc, int site)
{
    if (proc < 1 || proc > o->P || site < 1 || site > o->V) {
        o->status = ORC_BAD_ARG;
        return -1;
    }
    o->procstat[proc - 1] += 1;
    int c[4];
    nr2lattice(o, site, c);
    if (o->backend == 0) {
        const int32_t *pt = o->blob + HDR(o, 9) + 3 * (proc - 1);
        const int32_t *acts = o->blob + pt[2];
        for (int a = 0; a < pt[1]; ++a) {
            const int32_t *ac = acts + 5 * a;
            int t[4] = {c[0] + ac[1], c[1] + ac[2], c[2] + ac[3], ac[4]};
            if (routine_run(o, ac[0], t)) return -1;
        }
        return 0;
    }
    int g = o->blob[HDR(o, 9) + proc - 1];
    int cell[3] = {c[0], c[1], c[2]};
    return lat_int_run_group(o, g, cell);
}

/* One do_kmc_step (proclist_generic_subroutines.mpy:47-80). */
static int do_step(orc_t *o, int *proc_out, int *site_out)
{
    double u[3];
    update_accum_rate(o);
    if (!(o->accum[o->P - 1] > 0.0)) {
        o->status = ORC_ABSORBING;
        return -1;
    }
    orc_step_uniforms(o->seed, o->stream, o->rng_ctr, u);
    o->rng_ctr += 1;
    /* update_clocks(ran_time) */
    o->kmc_time_step = -log(u[0]) / o->accum[o->P - 1];
    o->kmc_time = o->kmc_time + o->kmc_time_step;
    o->kmc_step += 1;
    update_integ_rate(o);
    int proc, site;
    determine_procsite(o, u[1], u[2], &proc, &site);
    if (proc_out) *proc_out = proc;
    if (site_out) *site_out = site;
    return orc_run_proc_nr(o, proc, site);
}

/* do_kmc_steps(n); optional per-step trace of (proc, site, kmc_time).
 * Returns the number of steps executed (fewer than n if the replica
 * stopped; o->status says why). */
int64_t orc_run(orc_t *o, int64_t n, int32_t *tr_proc, int32_t *tr_site, double *tr_time)
{
    int64_t i;
    if (o->status != ORC_OK) return 0;
    for (i = 0; i < n; ++i) {
        int p = 0, s = 0;
        int rc = do_step(o, &p, &s);
        if (o->status == ORC_ABSORBING) break;
        if (tr_proc) tr_proc[i] = p;
        if (tr_site) tr_site[i] = s;
        if (tr_time) tr_time[i] = o->kmc_time;
        if (rc) { i += 1; break; }
    }
    return i;
}

/* ---------------- displacement tracking ---------------- */

/* initialize_mean_squared_displacement (base_acf.mpy:314-361): particles
 * numbered 1.. over the sites holding trace_species, in site order. */
void orc_msd_init(orc_t *o, int trace_species, const int32_t *hops, const double *disp)
{
    free(o->msd_hops); free(o->msd_d); free(o->msd_id); free(o->msd_site); free(o->msd_disp);
    o->msd_hops = malloc(sizeof(int32_t) * 9 * o->P);
    o->msd_d = malloc(sizeof(double) * 3 * o->P);
    memcpy(o->msd_hops, hops, sizeof(int32_t) * 9 * o->P);
    memcpy(o->msd_d, disp, sizeof(double) * 3 * o->P);
    o->msd_id = calloc(o->V + 1, sizeof(int32_t));
    o->msd_site = calloc(o->V + 1, sizeof(int32_t));
    o->msd_disp = calloc(3 * (size_t)(o->V + 1), sizeof(double));
    int i = 1;
    for (int k = 1; k <= o->V; ++k)
        if (o->lat[k] == trace_species) {
            o->msd_id[k] = i;
            o->msd_site[i] = k;
            i += 1;
        }
    o->msd_nions = i - 1;
}

/* allocate_trajectory (base_acf.mpy:427-458): events of the next n steps */
void orc_msd_trajectory(orc_t *o, int64_t n)
{
    free(o->msd_ev);
    o->msd_ev = n > 0 ? malloc(sizeof(int32_t) * 2 * (size_t)n) : NULL;
    o->msd_evcap = n;
    o->msd_evn = 0;
}

/* set_displacement_to_zero (base_acf.mpy:553-557) */
void orc_msd_zero(orc_t *o)
{
    memset(o->msd_disp, 0, sizeof(double) * 3 * (size_t)(o->V + 1));
}

void orc_msd_get(const orc_t *o, int32_t *nions, int32_t *site_arr, int32_t *id_arr, double *disp,
                 int64_t *nev, int32_t *ev)
{
    *nions = o->msd_nions;
    memcpy(site_arr, o->msd_site, sizeof(int32_t) * (o->V + 1));
    memcpy(id_arr, o->msd_id, sizeof(int32_t) * (o->V + 1));
    memcpy(disp, o->msd_disp, sizeof(double) * 3 * (size_t)(o->V + 1));
    *nev = o->msd_evn;
    if (ev && o->msd_evn)
        memcpy(ev, o->msd_ev, sizeof(int32_t) * 2 * (size_t)o->msd_evn);
}

/* do_kmc_steps_displacement (proclist_generic_subroutines_acf.mpy:143-219):
 * the step of do_step, with get_diff_sites_displacement (the hop table),
 * assign_particle_id, update_displacement and update_id_arr between
 * determine_procsite and run_proc_nr, and update_trajectory after it
 * (base_acf.mpy:825-891).  A process that is no hop moves no particle. */
int64_t orc_run_displacement(orc_t *o, int64_t n, int traj_on)
{
    int64_t i;
    if (o->status != ORC_OK || !o->msd_hops) return 0;
    for (i = 0; i < n; ++i) {
        double u[3];
        update_accum_rate(o);
        if (!(o->accum[o->P - 1] > 0.0)) {
            o->status = ORC_ABSORBING;
            break;
        }
        orc_step_uniforms(o->seed, o->stream, o->rng_ctr, u);
        o->rng_ctr += 1;
        o->kmc_time_step = -log(u[0]) / o->accum[o->P - 1];
        o->kmc_time = o->kmc_time + o->kmc_time_step;
        o->kmc_step += 1;
        update_integ_rate(o);
        int proc, site;
        determine_procsite(o, u[1], u[2], &proc, &site);
        const int32_t *h = o->msd_hops + 9 * (proc - 1);
        int particle = 0, fin = 0;
        if (h[0]) {
            int c[4];
            nr2lattice(o, site, c);
            int init = lattice2nr(o, c[0] + h[1], c[1] + h[2], c[2] + h[3], h[4]);
            fin = lattice2nr(o, c[0] + h[5], c[1] + h[6], c[2] + h[7], h[8]);
            particle = o->msd_id[init];                                  /* assign_particle_id */
            for (int d = 0; d < 3; ++d)                                  /* update_displacement */
                o->msd_disp[3 * particle + d] = o->msd_disp[3 * particle + d] + o->msd_d[3 * (proc - 1) + d];
            o->msd_id[init] = 0;                                         /* update_id_arr */
            o->msd_id[fin] = particle;
            o->msd_site[particle] = fin;
            o->msd_site[0] = 0;
        }
        int rc = orc_run_proc_nr(o, proc, site);
        if (traj_on && o->msd_evn < o->msd_evcap) {                      /* update_trajectory */
            o->msd_ev[2 * o->msd_evn] = particle;
            o->msd_ev[2 * o->msd_evn + 1] = fin;
            o->msd_evn += 1;
        }
        if (rc) { i += 1; break; }
    }
    return i;
}

/* ---------------- autocorrelation bins ---------------- */

/* allocate_tracing_arr + allocate_config_bin_acf (base_acf.mpy:382-524);
 * safety/extending 0 = the default factor 2 */
void orc_acf_alloc(orc_t *o, int nr_types, double t_bin, double t_f, int safety, int extending)
{
    if (safety == 0) safety = 2;
    if (extending == 0) extending = 2;
    free(o->acf_pa); free(o->acf_po); free(o->acf_buf); free(o->acf_pp); free(o->acf_cb); free(o->acf_cc);
    o->acf_nt1 = nr_types + 1;
    o->acf_pa = calloc(o->V + 1, sizeof(int32_t));
    o->acf_po = calloc(o->V + 1, sizeof(int32_t));
    o->acf_buf = calloc(o->V + 1, sizeof(double));
    o->acf_pp = calloc((size_t)o->acf_nt1 * o->acf_nt1, sizeof(double));
    if (!o->msd_id) {
        o->msd_id = calloc(o->V + 1, sizeof(int32_t));
        o->msd_site = calloc(o->V + 1, sizeof(int32_t));
        o->msd_disp = calloc(3 * (size_t)(o->V + 1), sizeof(double));
    }
    double tf = t_f * safety;
    o->acf_tbin = t_bin;
    o->acf_nbins = (int)(tf / t_bin + 1);
    o->acf_ext = o->acf_nbins * extending;
    o->acf_cb = calloc(o->acf_ext + 1, sizeof(double));     /* 1-based */
    o->acf_cc = calloc(o->acf_ext + 1, sizeof(int32_t));
    o->acf_bin = 0;
    o->acf_wrap = 0;
    o->acf_to = o->kmc_time;
}

/* set_property_acf for every site (pa[0..V]) and, if given, product_property */
void orc_acf_set(orc_t *o, const int32_t *pa, const double *pp)
{
    if (pa) memcpy(o->acf_pa, pa, sizeof(int32_t) * (o->V + 1));
    if (pp) memcpy(o->acf_pp, pp, sizeof(double) * o->acf_nt1 * o->acf_nt1);
}

/* update_property_and_buffer_acf (base_acf.mpy:1075-1096) */
static void acf_update_property_and_buffer(orc_t *o)
{
    for (int i = 1; i <= o->V; ++i) {
        o->acf_po[i] = o->acf_pa[o->msd_site[i]];
        o->acf_buf[i] = o->acf_pp[o->acf_po[i] * o->acf_nt1 + o->acf_po[i]];
    }
}

/* initialize_acf (base_acf.mpy:274-312) with the hop table of orc_msd_init */
void orc_acf_init(orc_t *o, int trace_species, const int32_t *hops)
{
    if (!o->msd_hops) {
        o->msd_hops = malloc(sizeof(int32_t) * 9 * o->P);
        o->msd_d = calloc(3 * (size_t)o->P, sizeof(double));
    }
    memcpy(o->msd_hops, hops, sizeof(int32_t) * 9 * o->P);
    int i = 1;
    for (int k = 1; k <= o->V; ++k) {
        if (o->lat[k] == trace_species) {
            o->msd_id[k] = i;
            o->msd_site[i] = k;
            i += 1;
        } else {
            o->msd_id[k] = 0;
        }
    }
    o->msd_nions = i - 1;
    o->acf_nr_before = o->msd_nions;
    o->msd_evn = 0;
    acf_update_property_and_buffer(o);
}

/* set_acf_to_zero (base_acf.mpy:526-551) */
void orc_acf_zero(orc_t *o)
{
    o->acf_nr_before = o->msd_nions;
    acf_update_property_and_buffer(o);
    memset(o->acf_cb, 0, sizeof(double) * (o->acf_ext + 1));
    memset(o->acf_cc, 0, sizeof(int32_t) * (o->acf_ext + 1));
    o->acf_bin = 0;
    o->acf_wrap = 0;
    o->acf_to = o->kmc_time;
}

/* update_config_bin (base_acf.mpy:1098-1172); the inner loop stops at the
 * last extended bin, where the reference would write past its array */
static void acf_update_config_bin(orc_t *o)
{
    double entry_bin = 0.0;
    for (int i = 0; i <= o->V; ++i)
        entry_bin = entry_bin + o->acf_buf[i];
    if (entry_bin == 0.0)
        entry_bin = 0.0;
    else
        entry_bin = entry_bin / (double)o->acf_nr_before;
    const double t = o->kmc_time, dt = o->kmc_time_step, t_bin = o->acf_tbin, t_o = o->acf_to;
    int b = o->acf_bin;
    if (t - t_o < (double)(b + 1) * t_bin) {
        o->acf_cb[b + 1] = o->acf_cb[b + 1] + entry_bin * dt;
        o->acf_cc[b + 1] += 1;
    } else {
        o->acf_cb[b + 1] = o->acf_cb[b + 1] + entry_bin * ((double)(b + 1) * t_bin - (t - t_o - dt));
        o->acf_cc[b + 1] += 1;
        b += 1;
        if (b < o->acf_nbins) {
            while ((double)(b + 1) * t_bin < t - t_o && b + 1 < o->acf_ext) {
                o->acf_cb[b + 1] = o->acf_cb[b + 1] + entry_bin * t_bin;
                o->acf_cc[b + 1] += 1;
                b += 1;
            }
            o->acf_cb[b + 1] = o->acf_cb[b + 1] + entry_bin * (t - t_o - (double)b * t_bin);
            o->acf_cc[b + 1] += 1;
        }
    }
    o->acf_bin = b;
}

/* do_kmc_steps_acf (proclist_generic_subroutines_acf.mpy:1-98) */
int64_t orc_run_acf(orc_t *o, int64_t n, int traj_on)
{
    int64_t i;
    if (o->status != ORC_OK || !o->msd_hops || !o->acf_cb) return 0;
    for (i = 0; i < n; ++i) {
        double u[3];
        update_accum_rate(o);
        if (!(o->accum[o->P - 1] > 0.0)) {
            o->status = ORC_ABSORBING;
            break;
        }
        orc_step_uniforms(o->seed, o->stream, o->rng_ctr, u);
        o->rng_ctr += 1;
        o->kmc_time_step = -log(u[0]) / o->accum[o->P - 1];
        o->kmc_time = o->kmc_time + o->kmc_time_step;
        o->kmc_step += 1;
        update_integ_rate(o);
        int proc, site;
        determine_procsite(o, u[1], u[2], &proc, &site);
        acf_update_config_bin(o);
        const int32_t *h = o->msd_hops + 9 * (proc - 1);
        int particle = 0, fin = 0;
        if (h[0]) {                                   /* get_diff_sites_acf */
            int c[4];
            nr2lattice(o, site, c);
            int init = lattice2nr(o, c[0] + h[1], c[1] + h[2], c[2] + h[3], h[4]);
            fin = lattice2nr(o, c[0] + h[5], c[1] + h[6], c[2] + h[7], h[8]);
            particle = o->msd_id[init];               /* assign_particle_id */
            o->msd_id[init] = 0;                      /* update_id_arr */
            o->msd_id[fin] = particle;
            o->msd_site[particle] = fin;
            o->msd_site[0] = 0;
        }
        if (o->acf_bin > o->acf_nbins - 1) {          /* update_after_wrap_acf */
            o->acf_to = o->kmc_time;
            o->acf_bin = 0;
            o->acf_wrap = o->acf_wrap + 1;
            acf_update_property_and_buffer(o);
            o->acf_nr_before = o->msd_nions;
        }
        int rc = orc_run_proc_nr(o, proc, site);
        /* update_buffer_acf */
        o->acf_buf[particle] = o->acf_pp[o->acf_po[particle] * o->acf_nt1 + o->acf_pa[o->msd_site[particle]]];
        if (traj_on && o->msd_evn < o->msd_evcap) {   /* update_trajectory */
            o->msd_ev[2 * o->msd_evn] = particle;
            o->msd_ev[2 * o->msd_evn + 1] = fin;
            o->msd_evn += 1;
        }
        if (rc) { i += 1; break; }
    }
    return i;
}

/* config_bin and counters (ext each), d4 = {sum(buffer_acf), t_o,
 * wrap_count, t_bin}, i5 = {bin_index, nr_of_bins, ext, nr_before, nions},
 * buffer_acf and property_o (V+1 each) */
void orc_acf_get(const orc_t *o, double *cb, int32_t *cc, double *d4, int32_t *i5, double *buf, int32_t *po)
{
    memcpy(cb, o->acf_cb + 1, sizeof(double) * o->acf_ext);
    memcpy(cc, o->acf_cc + 1, sizeof(int32_t) * o->acf_ext);
    double sum = 0.0;
    for (int i = 0; i <= o->V; ++i)
        sum = sum + o->acf_buf[i];
    d4[0] = sum; d4[1] = o->acf_to; d4[2] = o->acf_wrap; d4[3] = o->acf_tbin;
    i5[0] = o->acf_bin; i5[1] = o->acf_nbins; i5[2] = o->acf_ext; i5[3] = o->acf_nr_before; i5[4] = o->msd_nions;
    memcpy(buf, o->acf_buf, sizeof(double) * (o->V + 1));
    memcpy(po, o->acf_po, sizeof(int32_t) * (o->V + 1));
}

/* get_next_kmc_step (proclist_generic_subroutines.mpy:85-110): draws one
 * step's uniforms and passes ran_time as the site uniform. */
int orc_get_next_step(orc_t *o, int *proc, int *site)
{
    double u[3];
    update_accum_rate(o);
    if (!(o->accum[o->P - 1] > 0.0)) { o->status = ORC_ABSORBING; return -1; }
    orc_step_uniforms(o->seed, o->stream, o->rng_ctr, u);
    o->rng_ctr += 1;
    determine_procsite(o, u[1], u[0], proc, site);
    return 0;
}

/* proclist.get_occupation: [S][spuck] fractions, species-major */
void orc_get_occupation(const orc_t *o, double *occ)
{
    memset(occ, 0, sizeof(double) * o->S * o->spuck);
    for (int s = 1; s <= o->V; ++s) {
        int sp = o->lat[s];
        if (sp != o->nullsp && sp >= 0 && sp < o->S)
            occ[sp * o->spuck + ((s - 1) % o->spuck)] += 1.0;
    }
    double cells = (double)(o->X * o->Y * o->Z);
    for (int i = 0; i < o->S * o->spuck; ++i)
        occ[i] /= cells;
}

/* ---------------- accessors ---------------- */
int orc_status(const orc_t *o) { return o->status; }
int orc_volume(const orc_t *o) { return o->V; }
int orc_nr_of_proc(const orc_t *o) { return o->P; }
double orc_kmc_time(const orc_t *o) { return o->kmc_time; }
double orc_kmc_time_step(const orc_t *o) { return o->kmc_time_step; }
int64_t orc_kmc_step(const orc_t *o) { return o->kmc_step; }
uint64_t orc_rng_ctr(const orc_t *o) { return o->rng_ctr; }
uint64_t orc_touched(const orc_t *o) { return o->touched; }
void orc_reset_touched(orc_t *o) { o->touched = 0; }
void orc_set_kmc_time(orc_t *o, double t) { o->kmc_time = t; }
void orc_get_lattice(const orc_t *o, int8_t *out) { memcpy(out, o->lat + 1, o->V); }
void orc_get_nr(const orc_t *o, int32_t *out) { memcpy(out, o->nr, sizeof(int32_t) * o->P); }
void orc_get_procstat(const orc_t *o, int64_t *out) { memcpy(out, o->procstat, sizeof(int64_t) * o->P); }
void orc_get_integ(const orc_t *o, double *out) { memcpy(out, o->integ, sizeof(double) * o->P); }
void orc_get_accum(orc_t *o, double *out) { update_accum_rate(o); memcpy(out, o->accum, sizeof(double) * o->P); }
/* avail lists: sites [P][V] (slot order) and back [P][V] (site order) */
void orc_get_avail(const orc_t *o, int32_t *sites, int32_t *back)
{
    for (int p = 0; p < o->P; ++p) {
        memcpy(sites + (int64_t)p * o->V, o->sites + (int64_t)p * (o->V + 1) + 1, sizeof(int32_t) * o->V);
        memcpy(back + (int64_t)p * o->V, o->back + (int64_t)p * (o->V + 1) + 1, sizeof(int32_t) * o->V);
    }
}
/* reload_system (base.mpy:363-514): every field of the file overwrites the
 * state as it stands, without checks; integ_rates stay at zero unless given
 * (the file of save_system, base.mpy:517-578, has none). */
void orc_set_state(orc_t *o, double kmc_time, double kmc_time_step, int64_t kmc_step,
                   uint64_t rng_ctr, const int8_t *lattice, const int32_t *nr, const int64_t *procstat,
                   const double *rates, const double *integ, const int32_t *sites, const int32_t *back)
{
    o->kmc_time = kmc_time;
    o->kmc_time_step = kmc_time_step;
    o->kmc_step = kmc_step;
    o->rng_ctr = rng_ctr;
    o->status = 0;
    memcpy(o->lat + 1, lattice, o->V);
    memcpy(o->nr, nr, sizeof(int32_t) * o->P);
    memcpy(o->procstat, procstat, sizeof(int64_t) * o->P);
    memcpy(o->rates, rates, sizeof(double) * o->P);
    for (int p = 0; p < o->P; ++p) {
        o->integ[p] = integ ? integ[p] : 0.0;
        memcpy(o->sites + (int64_t)p * (o->V + 1) + 1, sites + (int64_t)p * o->V, sizeof(int32_t) * o->V);
        memcpy(o->back + (int64_t)p * (o->V + 1) + 1, back + (int64_t)p * o->V, sizeof(int32_t) * o->V);
    }
}
int orc_get_species(const orc_t *o, int site) { return o->lat[site]; }
int orc_replace_species(orc_t *o, int site, int old, int new_) { return replace_species(o, site, old, new_); }
