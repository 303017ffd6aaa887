/*
 * otf_oracle.c -- CPU restatement of the reference's otf step loop.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (and the cpu_baseline leg of
 * bench.py) through oracle/pyoracle.py, never by the engine.
 *
 * Follows kmos/fortran_src/base_otf.f90 and the otf proclist the reference
 * generates (kmos/io/__init__.py:2657-3723, fixtures under
 * tests/export_test/reference_export_otf and reference_export_intZGB_otf):
 *
 *   do_kmc_step (proclist.f90 of the fixtures): three uniforms ran_time,
 *     ran_proc, ran_site (kind rsingle, which kind_values.f90 defines as
 *     SELECTED_REAL_KIND(15, 200): double precision); update_accum_rate
 *     (every process's rate re-summed over its list, base_otf.f90:687-717);
 *     update_clocks (-log(ran_time) over the total, 1280-1321);
 *     update_integ_rate (719-739); determine_procsite (process
 *     by interval search over accum_rates, site by interval search over the
 *     running sum of the chosen process's rates, 1213-1278); run_proc_nr.
 *   del_proc / add_proc / update_rates_matrix: base_otf.f90:221-364.
 *   interval_search_real: base_otf.f90:1386-1490.
 *   initialize_state / touchup_cell / _adjust_database: the fixtures'
 *     proclist.f90 and kmos/run/__init__.py:1439-1457.
 *
 * The op programs, the bystander strides and the rate programs come from
 * the model blob (layout in include/kmcg.h).  The uniform stream is the
 * engine's Philox4x32-10 stream: ((v >> 12) + 0.5) * 2^-52 of a draw's
 * 64-bit word v.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { OTF_OK = 0, OTF_ABSORBING = 1, OTF_SPECIES_MISMATCH = 2, OTF_DEL_ABSENT = 3, OTF_BAD_ARG = 4,
       OTF_SEARCH = 5 };

typedef struct {
    int32_t *blob;
    int P, S, spuck, X, Y, Z, V, K;
    int8_t *lat;           /* lattice(1:V) */
    int32_t *a1, *a2;      /* avail_sites(p, i, 1) and (p, site, 2): [P][V+1] */
    double *rm;            /* rates_matrix(p, 1:V+1): [P][V+2] */
    int32_t *nsites;
    double *rates, *userpar, *table, *accum, *integ, *arr;
    int64_t *procstat;
    double kmc_time, kmc_time_step;
    int64_t kmc_step;
    uint32_t seed, stream;
    uint64_t ctr;
    int status;
    /* work counters (the bench's algorithmic bytes): list operations, list
     * entries whose running sum a step must re-add (from the first position
     * whose rate changed, in each list, on) and interval-search probes of a
     * binary search */
    int32_t *dmin;
    uint64_t n_ops, n_readd, n_probe;
} otf_t;

#define HDR(o, i) ((o)->blob[(i)])

/* ---- Philox4x32-10 and the uniforms ------------------------------------ */
static void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                   uint32_t out[4])
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u52(uint32_t lo, uint32_t hi)
{
    uint64_t v = ((uint64_t)hi << 32) | lo;
    return ((double)(v >> 12) + 0.5) * 0x1p-52;
}

static void step_uniforms(const otf_t *o, uint64_t step, double u[3])
{
    uint32_t w[4];
    philox((uint32_t)step, (uint32_t)(step >> 32), 0u, 0u, o->seed, o->stream, w);
    u[0] = u52(w[0], w[1]);
    u[1] = u52(w[2], w[3]);
    philox((uint32_t)step, (uint32_t)(step >> 32), 1u, 0u, o->seed, o->stream, w);
    u[2] = u52(w[0], w[1]);
}

/* ---- lattice ------------------------------------------------------------ */
static int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

static int nr_of(const otf_t *o, int x, int y, int z, int n)
{
    return o->spuck * (wrap(x, o->X) + o->X * (wrap(y, o->Y) + o->Y * wrap(z, o->Z))) + n;
}

static void cell_of(const otf_t *o, int site, int c[3])
{
    int k = (site - 1) / o->spuck;
    c[0] = k % o->X;
    c[1] = (k / o->X) % o->Y;
    c[2] = k / (o->X * o->Y);
}

/* ---- rate tables: the rate programs evaluated with Fortran kinds ------- */
typedef struct { int kind; int64_t i; float f; double d; } val_t;

static int64_t ipow_int(int64_t a, int64_t n)
{
    if (n < 0)
        return a == 1 ? 1 : (a == -1 ? ((n % 2 == 0) ? 1 : -1) : 0);
    int64_t r = 1;
    while (n--)
        r *= a;
    return r;
}

static double powi_d(double a, int64_t n)
{
    uint64_t u = n < 0 ? (uint64_t)(-n) : (uint64_t)n;
    double r = 1.0;
    while (u) {
        if (u & 1)
            r *= a;
        u >>= 1;
        if (u)
            a *= a;
    }
    return n < 0 ? 1.0 / r : r;
}

static float powi_f(float a, int64_t n)
{
    uint64_t u = n < 0 ? (uint64_t)(-n) : (uint64_t)n;
    float r = 1.0f;
    while (u) {
        if (u & 1)
            r *= a;
        u >>= 1;
        if (u)
            a *= a;
    }
    return n < 0 ? 1.0f / r : r;
}

static double eval_program(const int32_t *code, int ncode, const int *counts, double base,
                           const double *userpar, double *aux)
{
    val_t st[64];
    int sp = 0, i = 0;
    while (i < ncode) {
        int op = code[i];
        if (op == 0)
            break;
        switch (op) {
        case 1: st[sp].kind = 0; st[sp++].i = code[i + 1]; i += 2; continue;
        case 2: { float f; memcpy(&f, &code[i + 1], 4); st[sp].kind = 1; st[sp++].f = f; i += 2; continue; }
        case 3: st[sp].kind = 2; st[sp++].d = base; i += 1; continue;
        case 4: st[sp].kind = 2; st[sp++].d = userpar[code[i + 1]]; i += 2; continue;
        case 5: st[sp].kind = 0; st[sp++].i = counts[code[i + 1]]; i += 2; continue;
        case 6: st[sp].kind = 2; st[sp++].d = aux[code[i + 1]]; i += 2; continue;
        case 7: aux[code[i + 1]] = st[--sp].d; i += 2; continue;
        case 46: { double d; memcpy(&d, &code[i + 1], 8); st[sp].kind = 2; st[sp++].d = d; i += 3; continue; }
        default: break;
        }
        i += 1;
        val_t *a = &st[sp - 1];
        if (op == 8) { a->f = (float)a->i; a->kind = 1; }
        else if (op == 9) { a->d = (double)a->i; a->kind = 2; }
        else if (op == 10) { a->d = (double)a->f; a->kind = 2; }
        else if (op >= 11 && op <= 13) {
            if (a->kind == 0) a->i = -a->i;
            else if (a->kind == 1) a->f = -a->f;
            else a->d = -a->d;
        } else if (op >= 20 && op <= 31) {
            val_t *b = &st[sp - 1];
            a = &st[sp - 2];
            int k = (op - 20) % 3, base_op = op - k;
            if (k == 0) {
                if (base_op == 20) a->i = a->i + b->i;
                else if (base_op == 23) a->i = a->i - b->i;
                else if (base_op == 26) a->i = a->i * b->i;
                else a->i = b->i ? a->i / b->i : 0;        /* C truncates like Fortran */
            } else if (k == 1) {
                if (base_op == 20) a->f = a->f + b->f;
                else if (base_op == 23) a->f = a->f - b->f;
                else if (base_op == 26) a->f = a->f * b->f;
                else a->f = a->f / b->f;
            } else {
                if (base_op == 20) a->d = a->d + b->d;
                else if (base_op == 23) a->d = a->d - b->d;
                else if (base_op == 26) a->d = a->d * b->d;
                else a->d = a->d / b->d;
            }
            sp -= 1;
        } else if (op >= 32 && op <= 36) {
            val_t *b = &st[sp - 1];
            a = &st[sp - 2];
            if (op == 32) a->i = ipow_int(a->i, b->i);
            else if (op == 33) a->f = powi_f(a->f, b->i);
            else if (op == 34) a->d = powi_d(a->d, b->i);
            else if (op == 35) a->f = powf(a->f, b->f);
            else a->d = pow(a->d, b->d);
            sp -= 1;
        } else if (op >= 40 && op <= 45) {
            int dbl = (op - 40) % 2, fn = (op - 40) / 2;
            if (dbl) a->d = fn == 0 ? exp(a->d) : (fn == 1 ? log(a->d) : sqrt(a->d));
            else a->f = fn == 0 ? expf(a->f) : (fn == 1 ? logf(a->f) : sqrtf(a->f));
        } else {
            return NAN;
        }
    }
    return sp ? (st[sp - 1].kind == 2 ? st[sp - 1].d : NAN) : NAN;
}

static void build_tables(otf_t *o)
{
    const int rtab = HDR(o, 11);
    for (int q = 0; q < o->P; ++q) {
        const int32_t *rec = o->blob + rtab + 9 * q;
        int size = rec[0], tbase = rec[1], ncode = rec[4], naux = rec[6], nused = rec[7];
        const int32_t *code = o->blob + rec[5], *maxc = o->blob + rec[8];
        double aux[64];
        int counts[64];
        for (int t = 0; t < size; ++t) {
            int rem = t;
            for (int v = 0; v < nused; ++v) {
                counts[v] = rem % (maxc[v] + 1);
                rem /= maxc[v] + 1;
            }
            for (int a = 0; a < naux && a < 64; ++a)
                aux[a] = 0.0;
            o->table[tbase + t] = eval_program(code, ncode, counts, o->rates[q], o->userpar, aux);
        }
    }
}

/* gr_<q>(cell): the table entry of the bystander counts around the cell */
static double gr(const otf_t *o, int q, const int c[3])
{
    const int32_t *rec = o->blob + HDR(o, 11) + 9 * q;
    const int32_t *b = o->blob + rec[3];
    int idx = 0;
    for (int k = 0; k < rec[2]; ++k, b += 4 + o->S) {
        int sp = o->lat[nr_of(o, c[0] + b[0], c[1] + b[1], c[2] + b[2], b[3])];
        if (sp >= 0 && sp < o->S)
            idx += b[4 + sp];
    }
    return o->table[rec[1] + idx];
}

/* ---- base_otf.f90 list operations --------------------------------------- */
#define A1(o, p, i) ((o)->a1[(size_t)(p) * ((o)->V + 1) + (i)])
#define A2(o, p, s) ((o)->a2[(size_t)(p) * ((o)->V + 1) + (s)])
#define RM(o, p, i) ((o)->rm[(size_t)(p) * ((o)->V + 2) + (i)])

static void touch(otf_t *o, int p, int pos, int changed)
{
    o->n_ops += 1;
    if (changed && pos - 1 < o->dmin[p])
        o->dmin[p] = pos - 1;
}

static void del_proc(otf_t *o, int p, int site)
{
    int n = o->nsites[p];
    int mem = A2(o, p, site);
    if (mem == 0) {
        o->status = OTF_DEL_ABSENT;
        return;
    }
    touch(o, p, mem, mem < n && RM(o, p, mem) != RM(o, p, n));
    if (mem < n) {
        A1(o, p, mem) = A1(o, p, n);
        A1(o, p, n) = 0;
        RM(o, p, o->V + 1) = RM(o, p, o->V + 1) - RM(o, p, mem);
        RM(o, p, mem) = RM(o, p, n);
        RM(o, p, n) = 0.0;
        A2(o, p, A1(o, p, mem)) = mem;
    } else {
        A1(o, p, mem) = 0;
        RM(o, p, o->V + 1) = RM(o, p, o->V + 1) - RM(o, p, mem);
        RM(o, p, mem) = 0.0;
    }
    A2(o, p, site) = 0;
    o->nsites[p] = n - 1;
}

static void add_proc(otf_t *o, int p, int site, double rate)
{
    int n = ++o->nsites[p];
    touch(o, p, n, 1);
    A1(o, p, n) = site;
    A2(o, p, site) = n;
    RM(o, p, o->V + 1) = RM(o, p, o->V + 1) + rate;
    RM(o, p, n) = rate;
}

static void update_rates_matrix(otf_t *o, int p, int site, double rate)
{
    int mem = A2(o, p, site);
    touch(o, p, mem, RM(o, p, mem) != rate);
    RM(o, p, o->V + 1) = RM(o, p, o->V + 1) + rate - RM(o, p, mem);
    RM(o, p, mem) = rate;
}

/* the ops of one program at `cell` (0-based cell coordinates) */
static void run_ops(otf_t *o, const int32_t *w, int nops, const int c[3])
{
    for (int k = 0; k < nops; ++k) {
        int kind = w[0];
        if (kind == 1) {                                    /* replace_species */
            int site = nr_of(o, c[0] + w[1], c[1] + w[2], c[2] + w[3], w[4]);
            if (o->lat[site] != w[5])
                o->status = OTF_SPECIES_MISMATCH;
            o->lat[site] = (int8_t)w[6];
            w += 7;
            continue;
        }
        int q = w[1];
        int cc[3] = {c[0] + w[2], c[1] + w[3], c[2] + w[4]};
        int site = nr_of(o, cc[0], cc[1], cc[2], 1);
        int wc[3] = {wrap(cc[0], o->X), wrap(cc[1], o->Y), wrap(cc[2], o->Z)};
        if (kind == 0) {                                    /* if(can_do) del_proc */
            if (A2(o, q, site))
                del_proc(o, q, site);
            w += 5;
        } else if (kind == 2) {                             /* if(can_do) update_rates_matrix */
            if (A2(o, q, site))
                update_rates_matrix(o, q, site, gr(o, q, wc));
            w += 5;
        } else {                                            /* if-tree path, add_proc */
            int ncond = w[5], ok = 1;
            const int32_t *cd = w + 6;
            for (int j = 0; j < ncond && ok; ++j, cd += 5)
                ok = o->lat[nr_of(o, c[0] + cd[0], c[1] + cd[1], c[2] + cd[2], cd[3])] == cd[4];
            if (ok)
                add_proc(o, q, site, gr(o, q, wc));
            w += 6 + 5 * ncond;
        }
    }
}

static void touchup_cell(otf_t *o, int x, int y, int z)
{
    int site = nr_of(o, x, y, z, 1);
    for (int p = 0; p < o->P; ++p)
        if (A2(o, p, site))
            del_proc(o, p, site);
    int c[3] = {x, y, z};
    run_ops(o, o->blob + HDR(o, 12), HDR(o, 18), c);
}

/* ---- base_otf.f90:1386-1490 ---------------------------------------------- */
static int interval_search_real(otf_t *o, const double *arr, int n, double value)
{
    int left = 1, right = n, mid;
    for (;;) {
        mid = (right + left) >> 1;
        if (left >= right)
            break;
        if (value < arr[mid - 1])
            right = mid;
        else
            left = mid + 1;
    }
    if (arr[mid - 1] == 0.0) {
        for (;;) {
            if (arr[mid - 1] > 0.0) {
                if (mid >= n) {
                    o->status = OTF_SEARCH;
                    return mid;
                }
                break;
            }
            mid += 1;
            if (mid > n) {
                o->status = OTF_SEARCH;
                return n;
            }
        }
    }
    while (mid > 1 && arr[mid - 2] >= arr[mid - 1])
        mid -= 1;
    return mid;
}

/* ---- API ------------------------------------------------------------------ */
otf_t *otf_create(const int32_t *blob, int64_t nwords, int32_t X, int32_t Y, int32_t Z)
{
    if (nwords < 32 || blob[2] != 2)
        return NULL;
    otf_t *o = calloc(1, sizeof(otf_t));
    o->blob = malloc(sizeof(int32_t) * nwords);
    memcpy(o->blob, blob, sizeof(int32_t) * nwords);
    o->P = blob[3];
    o->S = blob[4];
    o->spuck = blob[6];
    o->X = X; o->Y = Y; o->Z = Z;
    o->V = X * Y * Z * o->spuck;
    o->K = blob[10];
    size_t P = o->P, V = o->V;
    o->lat = calloc(V + 1, 1);
    o->a1 = calloc(P * (V + 1), sizeof(int32_t));
    o->a2 = calloc(P * (V + 1), sizeof(int32_t));
    o->rm = calloc(P * (V + 2), sizeof(double));
    o->nsites = calloc(P, sizeof(int32_t));
    o->rates = calloc(P, sizeof(double));
    o->userpar = calloc(o->K + 1, sizeof(double));
    o->table = calloc(blob[13] + 1, sizeof(double));
    o->accum = calloc(P, sizeof(double));
    o->integ = calloc(P, sizeof(double));
    o->arr = calloc(V + 1, sizeof(double));
    o->procstat = calloc(P, sizeof(int64_t));
    o->dmin = calloc(P, sizeof(int32_t));
    return o;
}

void otf_free(otf_t *o)
{
    if (!o)
        return;
    free(o->blob); free(o->lat); free(o->a1); free(o->a2); free(o->rm); free(o->nsites);
    free(o->rates); free(o->userpar); free(o->table); free(o->accum); free(o->integ);
    free(o->arr); free(o->procstat); free(o->dmin); free(o);
}

/* initialize_state: every site its default species, touchup_cell over the
 * cells with x fastest; clocks and counters zero */
void otf_init(otf_t *o, uint32_t seed, uint32_t stream)
{
    const int32_t *def = o->blob + HDR(o, 8);
    size_t P = o->P, V = o->V;
    memset(o->a1, 0, sizeof(int32_t) * P * (V + 1));
    memset(o->a2, 0, sizeof(int32_t) * P * (V + 1));
    memset(o->rm, 0, sizeof(double) * P * (V + 2));
    memset(o->nsites, 0, sizeof(int32_t) * P);
    memset(o->accum, 0, sizeof(double) * P);
    memset(o->integ, 0, sizeof(double) * P);
    memset(o->procstat, 0, sizeof(int64_t) * P);
    for (int s = 1; s <= o->V; ++s)
        o->lat[s] = (int8_t)def[(s - 1) % o->spuck];
    for (int z = 0; z < o->Z; ++z)
        for (int y = 0; y < o->Y; ++y)
            for (int x = 0; x < o->X; ++x)
                touchup_cell(o, x, y, z);
    o->kmc_time = 0.0;
    o->kmc_time_step = 0.0;
    o->kmc_step = 0;
    o->seed = seed;
    o->stream = stream;
    o->ctr = 0;
    o->status = OTF_OK;
}

/* set_rate_const + update_user_parameter, then recalculate_rates_matrix:
 * every listed entry takes its cell's new rate, the list sums are
 * re-accumulated (reaccumulate_rates_matrix) */
void otf_set_params(otf_t *o, const double *rates, const double *userpar)
{
    memcpy(o->rates, rates, sizeof(double) * o->P);
    if (o->K)
        memcpy(o->userpar, userpar, sizeof(double) * o->K);
    build_tables(o);
    for (int p = 0; p < o->P; ++p) {
        for (int i = 1; i <= o->nsites[p]; ++i) {
            int c[3];
            cell_of(o, A1(o, p, i), c);
            RM(o, p, i) = gr(o, p, c);
        }
        RM(o, p, o->V + 1) = 0.0;
        for (int i = 1; i <= o->V; ++i) {
            if (A1(o, p, i) > 0)
                RM(o, p, o->V + 1) = RM(o, p, o->V + 1) + RM(o, p, i);
            else
                RM(o, p, i) = 0.0;
        }
    }
}

/* replace every site's species, then touchup_cell with x outermost
 * (_adjust_database, kmos/run/__init__.py:1439-1457) */
int otf_set_config(otf_t *o, const int8_t *config)
{
    for (int s = 1; s <= o->V; ++s) {
        if (config[s - 1] < 0 || config[s - 1] >= o->S)
            return OTF_BAD_ARG;
        o->lat[s] = config[s - 1];
    }
    for (int x = 0; x < o->X; ++x)
        for (int y = 0; y < o->Y; ++y)
            for (int z = 0; z < o->Z; ++z)
                touchup_cell(o, x, y, z);
    return o->status;
}

void otf_update_accum_rate(otf_t *o)
{
    for (int i = 0; i < o->P; ++i) {
        double s = 0.0;
        for (int j = 1; j <= o->nsites[i]; ++j)
            s = s + RM(o, i, j);
        RM(o, i, o->V + 1) = s;
        o->accum[i] = i == 0 ? s : o->accum[i - 1] + s;
    }
}

/* do_kmc_steps; with trace arrays, the chosen process (1-based), site and
 * kmc_time of every step */
int64_t otf_run(otf_t *o, int64_t n, int32_t *tproc, int32_t *tsite, double *ttime)
{
    int64_t done = 0;
    for (; done < n && o->status == OTF_OK; ++done) {
        double u[3];
        step_uniforms(o, o->ctr, u);
        otf_update_accum_rate(o);
        double total = o->accum[o->P - 1];
        if (!(total > 0.0)) {
            o->status = OTF_ABSORBING;
            break;
        }
        o->ctr += 1;
        o->kmc_time_step = -log(u[0]) / total;                      /* update_clocks */
        o->kmc_time = o->kmc_time + o->kmc_time_step;
        o->kmc_step += 1;
        for (int i = 0; i < o->P; ++i)                              /* update_integ_rate */
            o->integ[i] = o->integ[i] + RM(o, i, o->V + 1) * o->kmc_time_step;
        int proc = interval_search_real(o, o->accum, o->P, u[1] * total);
        int ns = o->nsites[proc - 1];
        if (ns <= 0) {
            o->status = OTF_SEARCH;
            break;
        }
        o->arr[0] = RM(o, proc - 1, 1);
        for (int i = 2; i <= ns; ++i)
            o->arr[i - 1] = o->arr[i - 2] + RM(o, proc - 1, i);
        int pos = interval_search_real(o, o->arr, ns, u[2] * o->arr[ns - 1]);
        int site = A1(o, proc - 1, pos);
        if (tproc) {
            tproc[done] = proc;
            tsite[done] = site;
            ttime[done] = o->kmc_time;
        }
        o->procstat[proc - 1] += 1;                                 /* run_proc_nr */
        int c[3];
        cell_of(o, site, c);
        for (int p = 0; p < o->P; ++p)
            o->dmin[p] = 0x7fffffff;
        for (int m = ns; m > 1; m = (m + 1) / 2)
            o->n_probe += 1;
        const int32_t *pt = o->blob + HDR(o, 9) + 2 * (proc - 1);
        run_ops(o, o->blob + pt[1], pt[0], c);
        for (int p = 0; p < o->P; ++p)
            if (o->dmin[p] < o->nsites[p])
                o->n_readd += (uint64_t)(o->nsites[p] - o->dmin[p]);
    }
    return done;
}

int otf_status(const otf_t *o) { return o->status; }
double otf_kmc_time(const otf_t *o) { return o->kmc_time; }
int64_t otf_kmc_step(const otf_t *o) { return o->kmc_step; }
uint64_t otf_rng_ctr(const otf_t *o) { return o->ctr; }
/* list operations, re-added running-sum entries, search probes so far */
void otf_work(const otf_t *o, uint64_t out[3])
{
    out[0] = o->n_ops;
    out[1] = o->n_readd;
    out[2] = o->n_probe;
}
int otf_volume(const otf_t *o) { return o->V; }
void otf_get_lattice(const otf_t *o, int8_t *out) { memcpy(out, o->lat + 1, o->V); }
void otf_get_nr(const otf_t *o, int32_t *out) { memcpy(out, o->nsites, sizeof(int32_t) * o->P); }
void otf_get_procstat(const otf_t *o, int64_t *out) { memcpy(out, o->procstat, sizeof(int64_t) * o->P); }
void otf_get_integ(const otf_t *o, double *out) { memcpy(out, o->integ, sizeof(double) * o->P); }
void otf_get_accum(const otf_t *o, double *out) { memcpy(out, o->accum, sizeof(double) * o->P); }
/* avail_sites(p, 1:V, 1:2) as [P][V][2] and rates_matrix(p, 1:V+1) as [P][V+1] */
void otf_get_avail(const otf_t *o, int32_t *avail, double *rm)
{
    for (int p = 0; p < o->P; ++p) {
        for (int i = 1; i <= o->V; ++i) {
            avail[((size_t)p * o->V + (i - 1)) * 2 + 0] = A1(o, p, i);
            avail[((size_t)p * o->V + (i - 1)) * 2 + 1] = A2(o, p, i);
        }
        for (int i = 1; i <= o->V + 1; ++i)
            rm[(size_t)p * (o->V + 1) + (i - 1)] = RM(o, p, i);
    }
}
void otf_get_table(const otf_t *o, double *out) { memcpy(out, o->table, sizeof(double) * o->blob[13]); }
/* continue the uniform stream at step `ctr` (the state of another run) */
void otf_set_clock(otf_t *o, double kmc_time, int64_t kmc_step, uint64_t ctr)
{
    o->kmc_time = kmc_time;
    o->kmc_step = kmc_step;
    o->ctr = ctr;
}
