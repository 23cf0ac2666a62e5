/*
 * gsnp_oracle.c -- CPU restatement of GSNP's SOAPsnp calling path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * anything under gsnp_amd/.
 *
 * Citations are reference file:line (paths relative to the reference root).
 *
 * Pins (the reference ships no golden vectors, so each function is pinned by
 * checks that a plausible mistake would fail; see tests/test_oracle_*.py):
 *   record codec ........ closed-form bit layout KATs, round trip
 *   sort ................ qsort vs brute force, 0xFFFFFFFF padding
 *   site matrix ......... per-read brute force vs the window builder;
 *                         invariants (sum count_uni = dep_uni = n, ...)
 *   score table ......... closed form per entry
 *   q_adj ............... closed form, -p 1 -g 1 gives q_adj = q
 *   likelihood .......... sorted-record form vs SOAPsnp's dense loop (two
 *                         independent statements, bit-equal); one record
 *                         gives the score row itself
 *   priors .............. hand KAT table (with and without -t, N rows)
 *   rank table .......... brute-force subset enumeration (N <= 14)
 *   normal_value ........ Abramowitz-Stegun 26.2.17 vs erfc, |err| < 7.5e-8
 *   posterior/q_cns ..... hand-computed sites covering each branch
 *   text ................ %#g string KATs, N-row literal
 *   -C codec ............ reference's own dictCompression (oracle/_ref,
 *                         built from cpp/dict_compression.h) and round trips
 */
#include "gsnp_oracle.h"

#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

static __thread char g_err[512]; /* per thread: the threaded run calls windows on several */

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char* orc_last_error(void) { return g_err; }

/* cpp/common.h:142-153 */
void orc_params_default(orc_params* p)
{
    memset(p, 0, sizeof *p);
    p->q_min = 64;
    p->q_max = 64 + 40;
    p->read_length = 45;
    p->althom_novel_r = 0.0005;
    p->het_novel_r = 0.0010;
    p->althom_val_r = 0.05;
    p->het_val_r = 0.10;
    p->althom_unval_r = 0.01;
    p->het_unval_r = 0.02;
    p->global_dependency = log10(0.9);
    p->pcr_dependency = log10(0.5);
}

static int check_params(const orc_params* p)
{
    if (p->read_length < 1 || p->read_length > 256)
        return fail(ORC_E_ARG, "read length %d outside [1,256]", p->read_length);
    if (p->q_min < 1 || p->q_max < p->q_min || p->q_max > 126 || p->q_max - p->q_min > 62)
        return fail(ORC_E_ARG, "quality chars q_min=%d q_max=%d invalid", p->q_min, p->q_max);
    return ORC_OK;
}

/* ============================================================== genome */

/* cpp/chromosome.cc:28-47 binarize: 4 bits per base, ((ch>>1)&7) */
static uint8_t bin_code(int ch) { return (uint8_t)((((unsigned)(signed char)ch) >> 1) & 7); }

static int read_line(FILE* f, char** buf, size_t* cap, size_t* len)
{
    size_t n = 0;
    int c;
    while ((c = fgetc(f)) != EOF) {
        if (n + 2 > *cap) {
            size_t nc = *cap ? *cap * 2 : 256;
            char* nb = (char*)realloc(*buf, nc);
            if (!nb) return -2;
            *buf = nb;
            *cap = nc;
        }
        if (c == '\n') break;
        (*buf)[n++] = (char)c;
    }
    if (c == EOF && n == 0) return -1;
    if (n + 1 > *cap) {
        char* nb = (char*)realloc(*buf, n + 1);
        if (!nb) return -2;
        *buf = nb;
        *cap = n + 1;
    }
    (*buf)[n] = 0;
    *len = n;
    return 0;
}

typedef struct {
    uint32_t pos;
    uint32_t order;
    orc_snp s;
} snp_tmp;

static int snp_tmp_cmp(const void* a, const void* b)
{
    const snp_tmp* x = (const snp_tmp*)a;
    const snp_tmp* y = (const snp_tmp*)b;
    if (x->pos != y->pos) return x->pos < y->pos ? -1 : 1;
    return x->order < y->order ? -1 : (x->order > y->order);
}

int orc_genome_find(const orc_genome* g, const char* name)
{
    for (int i = 0; i < g->n; i++)
        if (strcmp(g->chr[i].name, name) == 0) return i;
    return -1;
}

/* cpp/chromosome.cc:172-220 Genome::Genome (FASTA then dbSNP) */
int orc_genome_load(const char* fasta, const char* dbsnp, orc_genome** out)
{
    FILE* f = fopen(fasta, "r");
    if (!f) return fail(ORC_E_IO, "cannot open %s", fasta);
    orc_genome* g = (orc_genome*)calloc(1, sizeof *g);
    char* line = NULL;
    size_t cap = 0, len = 0;
    int cur = -1;
    size_t seq_cap = 0;
    int rc;
    while ((rc = read_line(f, &line, &cap, &len)) == 0) {
        if (len > 0 && line[0] == '>') {
            size_t i = 1;
            while (i < len && !isspace((unsigned char)line[i])) i++;
            char* name = (char*)malloc(i);
            memcpy(name, line + 1, i - 1);
            name[i - 1] = 0;
            if (orc_genome_find(g, name) >= 0) {
                free(name);
                fclose(f);
                free(line);
                orc_genome_free(g);
                return fail(ORC_E_FORMAT, "duplicate chromosome %s", name);
            }
            g->chr = (orc_chr*)realloc(g->chr, sizeof(orc_chr) * (g->n + 1));
            memset(&g->chr[g->n], 0, sizeof(orc_chr));
            g->chr[g->n].name = name;
            cur = g->n++;
            seq_cap = 0;
        } else if (cur >= 0) {
            orc_chr* c = &g->chr[cur];
            if (c->len + len + 1 > seq_cap) {
                seq_cap = (c->len + len + 1) * 2;
                c->code = (uint8_t*)realloc(c->code, seq_cap);
            }
            for (size_t i = 0; i < len; i++) c->code[c->len + i] = bin_code(line[i]);
            c->len += (uint32_t)len;
        }
    }
    fclose(f);
    if (rc == -2) return fail(ORC_E_NOMEM, "out of memory reading fasta");
    if (dbsnp && dbsnp[0]) {
        f = fopen(dbsnp, "r");
        if (!f) {
            free(line);
            orc_genome_free(g);
            return fail(ORC_E_IO, "cannot open %s", dbsnp);
        }
        /* per-chromosome temporary lists; first insertion wins (cpp/chromosome.cc:49-63) */
        snp_tmp** tmp = (snp_tmp**)calloc(g->n, sizeof(snp_tmp*));
        int* ntmp = (int*)calloc(g->n, sizeof(int));
        int* ctmp = (int*)calloc(g->n, sizeof(int));
        uint32_t order = 0;
        while ((rc = read_line(f, &line, &cap, &len)) == 0) {
            /* Format: Chr Pos hapmap validated indel fA fC fT fG rsID
               (cpp/chromosome.cc:205-217, cpp/soap_snp.h:735-738) */
            char name[256];
            long long pos;
            int hm, va, in;
            double fr[4];
            if (sscanf(line, "%255s %lld %d %d %d %lf %lf %lf %lf", name, &pos, &hm, &va, &in, &fr[0], &fr[1],
                       &fr[2], &fr[3]) != 9)
                continue;
            int ci = orc_genome_find(g, name);
            if (ci < 0) continue;
            if (pos < 1 || (uint64_t)pos > g->chr[ci].len) {
                fclose(f);
                free(line);
                return fail(ORC_E_RANGE, "dbSNP position %lld outside %s", pos, name);
            }
            if (ntmp[ci] == ctmp[ci]) {
                ctmp[ci] = ctmp[ci] ? ctmp[ci] * 2 : 64;
                tmp[ci] = (snp_tmp*)realloc(tmp[ci], sizeof(snp_tmp) * ctmp[ci]);
            }
            snp_tmp* t = &tmp[ci][ntmp[ci]++];
            t->pos = (uint32_t)(pos - 1);
            t->order = order++;
            t->s.hapmap = hm != 0;
            t->s.validated = va != 0;
            t->s.indel = in != 0;
            memcpy(t->s.freq, fr, sizeof fr);
        }
        fclose(f);
        for (int ci = 0; ci < g->n; ci++) {
            orc_chr* c = &g->chr[ci];
            if (!ntmp[ci]) continue;
            qsort(tmp[ci], ntmp[ci], sizeof(snp_tmp), snp_tmp_cmp);
            c->snp_pos = (uint32_t*)malloc(sizeof(uint32_t) * ntmp[ci]);
            c->snp = (orc_snp*)malloc(sizeof(orc_snp) * ntmp[ci]);
            int k = 0;
            for (int i = 0; i < ntmp[ci]; i++) {
                if (k && c->snp_pos[k - 1] == tmp[ci][i].pos) continue;
                c->snp_pos[k] = tmp[ci][i].pos;
                c->snp[k] = tmp[ci][i].s;
                c->code[tmp[ci][i].pos] |= 8; /* cpp/chromosome.cc:57 */
                k++;
            }
            c->nsnp = k;
            free(tmp[ci]);
        }
        free(tmp);
        free(ntmp);
        free(ctmp);
    }
    free(line);
    *out = g;
    return ORC_OK;
}

void orc_genome_free(orc_genome* g)
{
    if (!g) return;
    for (int i = 0; i < g->n; i++) {
        free(g->chr[i].name);
        free(g->chr[i].code);
        free(g->chr[i].snp_pos);
        free(g->chr[i].snp);
    }
    free(g->chr);
    free(g);
}

const orc_snp* orc_find_snp(const orc_chr* c, uint32_t pos)
{
    int lo = 0, hi = c->nsnp;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (c->snp_pos[mid] < pos)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < c->nsnp && c->snp_pos[lo] == pos) ? &c->snp[lo] : NULL;
}

/* =============================================================== reads */

/* istream extraction emulation for Soap_format::operator>> (cpp/soap_snp.h:132-169) */
typedef struct {
    const char* p;
} cursor;

static void skip_ws(cursor* c)
{
    while (*c->p && isspace((unsigned char)*c->p)) c->p++;
}

static int get_str(cursor* c, const char** s, int* n)
{
    skip_ws(c);
    if (!*c->p) return 0;
    *s = c->p;
    while (*c->p && !isspace((unsigned char)*c->p)) c->p++;
    *n = (int)(c->p - *s);
    return 1;
}

static int get_int(cursor* c, int* v)
{
    skip_ws(c);
    const char* q = c->p;
    int neg = 0;
    if (*q == '+' || *q == '-') neg = (*q++ == '-');
    if (!isdigit((unsigned char)*q)) return 0;
    long long x = 0;
    while (isdigit((unsigned char)*q)) {
        x = x * 10 + (*q++ - '0');
        if (x > 2147483648LL) return 0;
    }
    if (neg) x = -x;
    if (x > INT_MAX || x < INT_MIN) return 0;
    *v = (int)x;
    c->p = q;
    return 1;
}

static int get_char(cursor* c, char* ch)
{
    skip_ws(c);
    if (!*c->p) return 0;
    *ch = *c->p++;
    return 1;
}

static int is_acgt(int ch)
{
    switch (ch) {
    case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't':
        return 1;
    default:
        return 0;
    }
}

typedef struct {
    orc_reads* r;
    int64_t cap;
} reads_builder;

static void reads_grow(reads_builder* b)
{
    orc_reads* r = b->r;
    int64_t nc = b->cap ? b->cap * 2 : 1024;
    r->pos = (int32_t*)realloc(r->pos, sizeof(int32_t) * nc);
    r->len = (int32_t*)realloc(r->len, sizeof(int32_t) * nc);
    r->hit = (int32_t*)realloc(r->hit, sizeof(int32_t) * nc);
    r->chr = (int32_t*)realloc(r->chr, sizeof(int32_t) * nc);
    r->strand = (uint8_t*)realloc(r->strand, nc);
    r->seq = (char*)realloc(r->seq, (size_t)nc * r->L);
    r->qual = (char*)realloc(r->qual, (size_t)nc * r->L);
    b->cap = nc;
}

/* Parse all SOAP lines.  Lines that do not parse are skipped, as the
 * reference's `if (s >> soap)` does (cpp/matrix.cc:58-60).  Indel edits as
 * cpp/soap_snp.h:139-158; positions made 0-based (:160).  Input checks
 * (the explicit rejections this restatement makes instead of the
 * reference's undefined behaviour): read_len in [1, L]; hit in [0, 65535];
 * for unique hits every ACGT base with qual >= q_min has q in [1, 63]. */
int orc_reads_load(const char* soap, const orc_params* p, const orc_genome* g, orc_reads** out)
{
    int rc = check_params(p);
    if (rc) return rc;
    FILE* f = fopen(soap, "r");
    if (!f) return fail(ORC_E_IO, "cannot open %s", soap);
    reads_builder b = {(orc_reads*)calloc(1, sizeof(orc_reads)), 0};
    b.r->L = p->read_length;
    const int L = p->read_length;
    char* line = NULL;
    size_t cap = 0, len = 0;
    char* work = (char*)malloc(4 * 1024);
    size_t work_cap = 4 * 1024;
    while ((rc = read_line(f, &line, &cap, &len)) == 0) {
        cursor c = {line};
        const char *id, *rd, *ql, *chr;
        int nid, nrd, nql, nchr, hit, rlen, pos, mm;
        char ab, strand;
        if (!get_str(&c, &id, &nid) || !get_str(&c, &rd, &nrd) || !get_str(&c, &ql, &nql) || !get_int(&c, &hit) ||
            !get_char(&c, &ab) || !get_int(&c, &rlen) || !get_char(&c, &strand) || !get_str(&c, &chr, &nchr) ||
            !get_int(&c, &pos) || !get_int(&c, &mm))
            continue;
        /* edited read/qual strings */
        size_t need = (size_t)(nrd + nql) + 2u * 1024 + 64;
        if (need > work_cap) {
            work_cap = need * 2;
            work = (char*)realloc(work, work_cap);
        }
        const char* eread = rd;
        const char* equal = ql;
        int elen = nrd, eqlen = nql;
        if (mm > 100) {
            int ip;
            if (!get_int(&c, &ip)) continue;
            char* wr = work;
            char* wq = work + work_cap / 2;
            if (mm > 200) {
                int il = mm - 200;
                if (ip < 0 || ip > nrd || ip > nql || il > 1024) {
                    fclose(f);
                    free(line);
                    free(work);
                    orc_reads_free(b.r);
                    return fail(ORC_E_FORMAT, "indel position out of range");
                }
                /* read.substr(0,ip) + N*il + read.substr(ip, rlen-ip) */
                int tail = rlen - ip;
                if (tail < 0) tail = 0;
                int trd = nrd - ip < tail ? nrd - ip : tail;
                int tql = nql - ip < tail ? nql - ip : tail;
                memcpy(wr, rd, ip);
                memset(wr + ip, 'N', il);
                memcpy(wr + ip + il, rd + ip, trd);
                elen = ip + il + trd;
                memcpy(wq, ql, ip);
                memset(wq + ip, 'N', il);
                memcpy(wq + ip + il, ql + ip, tql);
                eqlen = ip + il + tql;
            } else {
                int il = mm - 100;
                if (ip < 0 || ip + il > nrd || ip + il > nql) {
                    fclose(f);
                    free(line);
                    free(work);
                    orc_reads_free(b.r);
                    return fail(ORC_E_FORMAT, "indel position out of range");
                }
                int tail = rlen - ip - il;
                if (tail < 0) tail = 0;
                int trd = nrd - ip - il < tail ? nrd - ip - il : tail;
                int tql = nql - ip - il < tail ? nql - ip - il : tail;
                memcpy(wr, rd, ip);
                memcpy(wr + ip, rd + ip + il, trd);
                elen = ip + trd;
                memcpy(wq, ql, ip);
                memcpy(wq + ip, ql + ip + il, tql);
                eqlen = ip + tql;
            }
            eread = wr;
            equal = wq;
        }
        pos -= 1;
        int ci = -1;
        if (pos >= 0) {
            char nm[1024];
            int k = nchr < 1023 ? nchr : 1023;
            memcpy(nm, chr, k);
            nm[k] = 0;
            ci = orc_genome_find(g, nm);
            if (ci < 0) {
                fclose(f);
                free(line);
                free(work);
                orc_reads_free(b.r);
                return fail(ORC_E_CHROM, "Chromosome: !%s! NOT found", nm);
            }
            if (rlen < 1 || rlen > L) {
                fclose(f);
                free(line);
                free(work);
                orc_reads_free(b.r);
                return fail(ORC_E_RANGE, "read length %d outside [1,%d]", rlen, L);
            }
            if (hit < 0 || hit > 65535) {
                fclose(f);
                free(line);
                free(work);
                orc_reads_free(b.r);
                return fail(ORC_E_RANGE, "hit count %d outside [0,65535]", hit);
            }
        }
        if (b.r->n == b.cap) reads_grow(&b);
        int64_t i = b.r->n++;
        b.r->pos[i] = pos;
        b.r->len[i] = rlen;
        b.r->hit[i] = hit;
        b.r->chr[i] = ci;
        b.r->strand[i] = (uint8_t)strand;
        char* s = b.r->seq + (size_t)i * L;
        char* q = b.r->qual + (size_t)i * L;
        for (int k = 0; k < L; k++) {
            s[k] = k < elen ? eread[k] : 'N';
            q[k] = k < eqlen ? equal[k] : 0;
        }
        if (pos >= 0 && hit == 1) {
            for (int k = 0; k < rlen; k++) {
                if (!is_acgt(s[k])) continue;
                int qc = (unsigned char)q[k];
                if (qc < p->q_min) continue;
                int qq = qc - p->q_min;
                if (qq < 1 || qq > 63) {
                    fclose(f);
                    free(line);
                    free(work);
                    orc_reads_free(b.r);
                    return fail(ORC_E_QUAL, "quality %d outside [1,63] on a unique-hit base", qq);
                }
            }
        }
    }
    fclose(f);
    free(line);
    free(work);
    *out = b.r;
    return ORC_OK;
}

void orc_reads_free(orc_reads* r)
{
    if (!r) return;
    free(r->pos);
    free(r->len);
    free(r->hit);
    free(r->chr);
    free(r->strand);
    free(r->seq);
    free(r->qual);
    free(r);
}

/* ============================================================== tables */

/* cpp/matrix.cc:42-213 Prob_matrix::matrix_gen (p_matrix starts at 1.0, :11-13) */
int orc_matrix_gen(const orc_reads* r, const orc_genome* g, const orc_params* p, double* pm)
{
    int rc = check_params(p);
    if (rc) return rc;
    uint64_t* cnt = (uint64_t*)calloc(ORC_PMATRIX, sizeof(uint64_t));
    if (!cnt) return fail(ORC_E_NOMEM, "count matrix");
    for (int i = 0; i < ORC_PMATRIX; i++) pm[i] = 1.0;
    const int L = r->L;
    for (int64_t i = 0; i < r->n; i++) {
        if (r->pos[i] < 0) continue;                                   /* :65 */
        const orc_chr* c = &g->chr[r->chr[i]];
        if ((int64_t)r->pos[i] + r->len[i] >= (int64_t)c->len) continue; /* :83 */
        if (r->hit[i] != 1) continue;                                  /* :86 */
        const char* s = r->seq + (size_t)i * L;
        const char* q = r->qual + (size_t)i * L;
        for (int coord = 0; coord < r->len[i]; coord++) {
            if (!is_acgt(s[coord])) continue; /* :88 'N' (other chars abort in the reference's codec) */
            uint64_t ref = c->code[r->pos[i] + coord];
            if ((ref & 12) != 0) continue; /* :100 ref N or dbSNP */
            uint64_t cyc = r->strand[i] == '+' ? (uint64_t)coord : (uint64_t)(r->len[i] - 1 - coord);
            uint64_t qc = (unsigned char)q[coord];
            cnt[(qc << 12) | (cyc << 4) | ((ref & 3) << 2) | (((unsigned)s[coord] >> 1) & 3)] += 1;
        }
    }
    const uint64_t sta_pow = 10; /* :130 */
    for (int q_char = p->q_min; q_char <= p->q_max; q_char++) {
        uint64_t by_type[16] = {0}, by_t[4] = {0}, total = 0, mism = 0;
        for (int coord = 0; coord != p->read_length; coord++) {
            for (int type = 0; type != 16; type++) {
                uint64_t v = cnt[((uint64_t)q_char << 12) | ((uint64_t)coord << 4) | type];
                by_type[type] += v;
                by_t[(type >> 2) & 3] += v;
                total += v;
                if (type % 5 != 0) mism += v;
            }
        }
        for (int coord = 0; coord != p->read_length; coord++) {
            uint64_t sum[4] = {0};
            for (int type = 0; type != 16; type++)
                sum[(type >> 2) & 3] += cnt[((uint64_t)q_char << 12) | ((uint64_t)coord << 4) | type];
            for (int t = 0; t != 4; t++) {
                for (int o = 0; o != 4; o++) {
                    uint64_t ci = ((uint64_t)q_char << 12) | ((uint64_t)coord << 4) | (t << 2) | o;
                    uint64_t pi = ((uint64_t)(q_char - p->q_min) << 12) | ((uint64_t)coord << 4) | (t << 2) | o;
                    if (cnt[ci] > sta_pow)
                        pm[pi] = ((double)cnt[ci]) / (double)sum[t];
                    else if (by_type[t << 2 | o] > sta_pow)
                        pm[pi] = ((double)by_type[t << 2 | o]) / (double)by_t[t];
                    else if (total > 0) {
                        if (o == t)
                            pm[pi] = ((double)(total - mism)) / (double)total;
                        else
                            pm[pi] = ((double)mism) / (double)total;
                    }
                    if (pm[pi] == 0 || pm[pi] == 1) { /* :187-199 */
                        if (o == t) {
                            pm[pi] = (1 - pow(10, -((q_char - p->q_min) / 10.0)));
                            if (pm[pi] < 0.25) pm[pi] = 0.25;
                        } else {
                            pm[pi] = (pow(10, -((q_char - p->q_min) / 10.0)) / 3);
                            if (pm[pi] > 0.25) pm[pi] = 0.25;
                        }
                    }
                }
            }
        }
    }
    free(cnt);
    return ORC_OK;
}

/* cpp/matrix.cc:230-241: "q\tcoord" then 16 x "\t%.16e" */
int orc_matrix_write(const char* path, const double* pm, const orc_params* p)
{
    FILE* f = fopen(path, "w");
    if (!f) return fail(ORC_E_IO, "cannot create %s", path);
    for (int q = 0; q <= p->q_max - p->q_min; q++) {
        for (int coord = 0; coord != p->read_length; coord++) {
            fprintf(f, "%d\t%d", q, coord);
            for (int t = 0; t != 16; t++) fprintf(f, "\t%.16e", pm[((uint64_t)q << 12) | ((uint64_t)coord << 4) | t]);
            fputc('\n', f);
        }
    }
    fclose(f);
    return ORC_OK;
}

/* cpp/matrix.cc:215-228 */
int orc_matrix_read(const char* path, double* pm)
{
    FILE* f = fopen(path, "r");
    if (!f) return fail(ORC_E_IO, "cannot open %s", path);
    for (int i = 0; i < ORC_PMATRIX; i++) pm[i] = 1.0;
    char* line = NULL;
    size_t cap = 0, len = 0;
    while (read_line(f, &line, &cap, &len) == 0) {
        char* s = line;
        char* e;
        long q = strtol(s, &e, 10);
        if (e == s) continue;
        s = e;
        long coord = strtol(s, &e, 10);
        if (e == s) continue;
        s = e;
        if (q < 0 || q > 255 || coord < 0 || coord > 255) continue;
        for (int t = 0; t != 16; t++) {
            double v = strtod(s, &e);
            if (e == s) break;
            s = e;
            pm[((uint64_t)q << 12) | ((uint64_t)coord << 4) | t] = v;
        }
    }
    free(line);
    fclose(f);
    return ORC_OK;
}

/* cpp/prior.cc:3-43 Prob_matrix::prior_gen (p_prior starts at 1.0, cpp/matrix.cc:14-16) */
void orc_prior_gen(const orc_params* p, double* pp)
{
    for (int i = 0; i < 128; i++) pp[i] = 1.0;
    for (int t = 0; t != 4; t++) {
        for (int a1 = 0; a1 != 4; a1++) {
            for (int a2 = a1; a2 != 4; a2++) {
                int k = t << 4 | a1 << 2 | a2;
                if (a1 == t && a2 == t)
                    pp[k] = 1;
                else if (a1 == t || a2 == t)
                    pp[k] = p->het_novel_r;
                else if (a1 == a2)
                    pp[k] = p->althom_novel_r;
                else
                    pp[k] = p->het_novel_r * p->althom_novel_r;
                if (p->transition_dominant && ((a1 ^ t) == 0x3 || (a2 ^ t) == 0x3)) pp[k] *= 4;
            }
        }
    }
    for (int a1 = 0; a1 != 4; a1++) {
        for (int a2 = a1; a2 != 4; a2++) {
            for (int t = 4; t < 8; t++)
                pp[t << 4 | a1 << 2 | a2] = (a1 == a2 ? 1 : (2 * p->het_novel_r)) * 0.25 * 0.25;
        }
    }
}

/* cpp/rank_sum.cc:3-42 Prob_matrix::rank_table_gen (p_rank starts at 1.0, cpp/matrix.cc:24-26) */
void orc_rank_table_gen(double* p_rank)
{
    double fact[64];
    fact[0] = 1.0;
    for (int i = 1; i != 64; i++) fact[i] = fact[i - 1] * i;
    for (int i = 0; i < ORC_PRANK; i++) p_rank[i] = 1.0;
    uint64_t* rs = (uint64_t*)calloc(ORC_PRANK, sizeof(uint64_t));
    rs[0] = 1;
    for (uint64_t N = 1; N != 64; N++) {
        for (uint64_t n1 = 0; n1 <= N; n1++) {
            for (uint64_t T1 = (1 + n1) * n1 / 2; T1 <= (N + N - n1 + 1) * n1 / 2; T1++) {
                rs[N << 17 | n1 << 11 | T1] = rs[((N - 1) << 17) | (n1 << 11) | T1] +
                                              ((T1 >= N && n1 > 0) ? rs[((N - 1) << 17) | ((n1 - 1) << 11) | (T1 - N)] : 0);
                p_rank[(N << 17) | (n1 << 11) | T1] = rs[N << 17 | n1 << 11 | T1] / (fact[N] / (fact[n1] * fact[N - n1]));
            }
            double p_left = 0.0, p_right = 1.0;
            for (uint64_t T1 = (1 + n1) * n1 / 2; T1 <= (N + N - n1 + 1) * n1 / 2; T1++) {
                p_right = 1.0 - p_left;
                p_left += p_rank[(N << 17) | (n1 << 11) | T1];
                p_rank[N << 17 | n1 << 11 | T1] = (p_left < p_right ? p_left : p_right);
            }
        }
    }
    free(rs);
}

/* kernel/likelihood_kernel.cu:137-161 materializePMatrix */
void orc_score_table(const double* pm, double* S)
{
    for (int q = 0; q < 256; q++)
        for (int coord = 0; coord < 256; coord++)
            for (int o = 0; o < 4; o++) {
                uint32_t base = ((uint32_t)q << 10) | (coord << 2) | o;
                int k = 0;
                for (int a1 = 0; a1 < 4; a1++)
                    for (int a2 = a1; a2 < 4; a2++) {
                        /* cpp/gsnp.h:16 P_MAT_IDX */
                        double x = pm[((uint64_t)q << 12) | (coord << 4) | (a1 << 2) | o];
                        double y = pm[((uint64_t)q << 12) | (coord << 4) | (a2 << 2) | o];
                        S[(size_t)base * 10 + k] = log10(0.5 * x + 0.5 * y);
                        k++;
                    }
            }
}

/* ====================================================== site primitives */

/* kernel/likelihood_kernel.cu:24-42 */
uint32_t orc_encode_word(uint32_t coord, uint32_t strand, uint32_t o_base, uint32_t q, uint32_t occ)
{
    return (o_base << 23) | ((64 - q) << 17) | (coord << 9) | (strand << 8) | occ;
}

void orc_decode_word(uint32_t w, uint32_t* coord, uint32_t* strand, uint32_t* o_base, uint32_t* q, uint32_t* occ)
{
    *o_base = (w & 0x01800000) >> 23;
    *q = 64 - ((w & 0x007e0000) >> 17);
    *coord = (w & 0x0001fe00) >> 9;
    *strand = (w & 0x00000100) >> 8;
    *occ = w & 0xff;
}

static int u32_cmp(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y);
}

/* kernel/batchsort.cu:162-291 + kernel/bitonic_kernel.cu:119-176: ordering contract only
 * (each site's records ascending as u32) */
void orc_sort_records(uint32_t* rec, int n) { qsort(rec, n, sizeof(uint32_t), u32_cmp); }

/* kernel/likelihood_kernel.cu:226-231 */
int orc_qadj(int q, int t, int g, double pcr, double glob)
{
    double l = log10((double)q); /* kernel/likelihood_kernel.cu:83-87 d_log10Table */
    int qa = (int)(pow(10.0, (l + (t - 1) * pcr + g * glob)) + 0.5);
    if (qa < 1) qa = 1;
    return qa;
}

/* kernel/likelihood_kernel.cu:168-256 likelihoodSparse_kernel, one site, records sorted.
 * lk10 is indexed by the 10 genotypes a1<=a2 in the order AA AC AT AG CC CT CG TT TG GG. */
int orc_likelihood_sorted(const uint32_t* rec, int n, int read_length, double pcr, double glob, const double* S,
                          double* lk10)
{
    int cnt[512];
    int last_o = -1, gdc = -1;
    for (int k = 0; k < 10; k++) lk10[k] = 0.0;
    for (int i = 0; i < n; i++) {
        uint32_t coord, strand, o, q, occ;
        orc_decode_word(rec[i], &coord, &strand, &o, &q, &occ);
        if ((int)o != last_o) {
            gdc = -1;
            memset(cnt, 0, sizeof cnt);
            last_o = (int)o;
        }
        int idx = (int)(strand * read_length + coord);
        int tmp = cnt[idx];
        for (uint32_t k = 0; k < occ; k++) {
            if (tmp == 0) gdc += 1;
            tmp++;
            int qa = orc_qadj((int)q, tmp, gdc, pcr, glob);
            if (qa > 255) return fail(ORC_E_RANGE, "adjusted quality %d exceeds the score table", qa);
            const double* row = S + (size_t)((((uint32_t)qa) << 10) | (coord << 2) | o) * 10;
            for (int g = 0; g < 10; g++) lk10[g] += row[g];
        }
        cnt[idx] = tmp;
    }
    return ORC_OK;
}

/* The same sum stated SOAPsnp's dense way: counts in a base x strand x q x cycle
 * matrix, loops o_base -> q descending -> cycle ascending -> strand (the original
 * SOAPsnp likelihood loop that GSNP's sorted records replace; the loop nest is
 * the one rank_test still uses, cpp/rank_sum.cc:85-97).  Independent of the
 * sort: checks that ascending-u32 order equals this nest order. */
int orc_likelihood_dense(const uint32_t* rec, int n, int read_length, int q_max_min, double pcr, double glob,
                         const double* S, double* lk10)
{
    static uint8_t info[4 * 2 * 64 * 256];
    memset(info, 0, sizeof info);
    for (int i = 0; i < n; i++) {
        uint32_t coord, strand, o, q, occ;
        orc_decode_word(rec[i], &coord, &strand, &o, &q, &occ);
        info[o << 15 | strand << 14 | (q & 63) << 8 | coord] += (uint8_t)occ;
    }
    for (int k = 0; k < 10; k++) lk10[k] = 0.0;
    int cnt[512];
    for (int o = 0; o < 4; o++) {
        int gdc = -1;
        memset(cnt, 0, sizeof cnt);
        for (int q = q_max_min; q >= 0; q--) {
            for (int coord = 0; coord < read_length; coord++) {
                for (int strand = 0; strand < 2; strand++) {
                    int occ = info[o << 15 | strand << 14 | q << 8 | coord];
                    int idx = strand * read_length + coord;
                    for (int k = 0; k < occ; k++) {
                        if (cnt[idx] == 0) gdc += 1;
                        cnt[idx]++;
                        int qa = orc_qadj(q, cnt[idx], gdc, pcr, glob);
                        if (qa > 255) return fail(ORC_E_RANGE, "adjusted quality exceeds the score table");
                        const double* row = S + (size_t)((((uint32_t)qa) << 10) | (coord << 2) | o) * 10;
                        for (int g = 0; g < 10; g++) lk10[g] += row[g];
                    }
                }
            }
        }
    }
    return ORC_OK;
}

/* cpp/normal_dis.cc:3-24 (Abramowitz & Stegun 26.2.17) */
double orc_normal_value(double z)
{
    if (z > 6.0 || z < -6.0) return 0.0;
    double b1 = 0.31938153, b2 = -0.356563782, b3 = 1.781477937, b4 = -1.821255978, b5 = 1.330274429;
    double pp = 0.2316419, c2 = 0.39894228;
    double a = fabs(z);
    double t = 1.0 / (1.0 + a * pp);
    double b = c2 * exp((-z) * (z / 2.0));
    double nn = ((((b5 * t + b4) * t + b3) * t + b2) * t + b1) * t;
    nn = 1.0 - b * nn;
    if (z < 0.0) nn = 1.0 - nn;
    return nn > 0.5 ? 1 - nn : nn;
}

/* cpp/rank_sum.cc:44-49 (n1*(n1+n2+1)/2 is integer division) */
double orc_normal_test(int n1, int n2, double T1, double T2)
{
    double u1 = (T1 - n1 * (n1 + n2 + 1) / 2) / sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    double u2 = (T2 - n2 * (n1 + n2 + 1) / 2) / sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    return orc_normal_value(fabs(u1) > fabs(u2) ? u1 : u2);
}

/* cpp/rank_sum.cc:51-58, including the <<16 interpolation index */
double orc_table_test(const double* p_rank, int n1, int n2, double T1, double T2)
{
    if (n1 <= n2)
        return p_rank[(n1 + n2) << 17 | n1 << 11 | (int)(T1)] +
               (T1 - (int)T1) * (p_rank[(n1 + n2) << 16 | n1 << 11 | (int)(T1 + 1)] -
                                 p_rank[(n1 + n2) << 17 | n1 << 11 | (int)(T1)]);
    return p_rank[(n1 + n2) << 17 | n2 << 11 | (int)(T2)] +
           (T2 - (int)T2) * (p_rank[(n1 + n2) << 16 | n2 << 11 | (int)(T2 + 1)] -
                             p_rank[(n1 + n2) << 17 | n2 << 11 | (int)(T2)]);
}

/* kernel/posterior_kernel.cu:506-572 gpu_rank_test over sorted records */
double orc_rank_test(const uint32_t* rec, int n, const int* cu, int q_max_min, int best_type, const double* p_rank)
{
    int a = best_type & 3, b = (best_type >> 2) & 3;
    if (a == b) return 1.0;
    if (cu[a] == 0 || cu[b] == 0) return 0.0;
    int same[64];
    double rank_array[64];
    double T[4] = {0, 0, 0, 0};
    int need[4] = {0, 0, 0, 0};
    memset(same, 0, sizeof same);
    for (int i = 0; i < 64; i++) rank_array[i] = 0.0;
    need[a] = need[b] = 1;
    for (int i = 0; i < n; i++) {
        int o = (rec[i] & 0x01800000) >> 23;
        int q = 64 - ((rec[i] & 0x007e0000) >> 17);
        int occ = rec[i] & 0xff;
        if (cu[o] == 0 || !need[o]) continue;
        same[q & 63] += occ; /* q<=63 after input checks */
    }
    int rank = 0;
    for (int q = 0; q <= q_max_min + 1 && q < 64; q++) {
        rank_array[q] = rank + (1 + same[q]) / 2.0;
        rank += same[q];
    }
    for (int i = 0; i < n; i++) {
        int o = (rec[i] & 0x01800000) >> 23;
        int q = 64 - ((rec[i] & 0x007e0000) >> 17);
        int occ = rec[i] & 0xff;
        if (cu[o] == 0 || !need[o]) continue;
        T[o] += rank_array[q & 63] * occ;
    }
    if (cu[a] + cu[b] < 64) return orc_table_test(p_rank, cu[a], cu[b], T[a], T[b]);
    return orc_normal_test(cu[a], cu[b], T[a], T[b]);
}

/* kernel/posterior_kernel.cu:575-633 gpu_snp_p_prior_gen (the device path, which
 * has no allele-count exit) */
int orc_snp_prior(double* pr, int ref, const orc_snp* s, const orc_params* p)
{
    if (s->indel) return 0;
    int t = ref & 3;
    for (int a1 = 0; a1 != 4; a1++) {
        for (int a2 = a1; a2 != 4; a2++) {
            if (!s->hapmap) {
                if (s->freq[a1] > 0 && s->freq[a2] > 0) {
                    if (a1 == a2 && a1 == t)
                        pr[a1 << 2 | a2] = 1;
                    else if (a1 == t || a2 == t)
                        pr[a1 << 2 | a2] = s->validated ? p->het_val_r : p->het_unval_r;
                    else if (a1 == a2)
                        pr[a1 << 2 | a2] = s->validated ? p->althom_val_r : p->althom_unval_r;
                    else
                        pr[a1 << 2 | a2] = s->validated ? p->het_val_r : p->het_unval_r;
                }
            } else if (s->freq[a1] > 0 && s->freq[a2] > 0) {
                pr[a1 << 2 | a2] = (a1 == a2 ? 1 : (2 * p->het_val_r)) * s->freq[a1] * s->freq[a2];
            }
        }
    }
    return 1;
}

/* (int)x as x86-64 cvttsd2si: NaN and out-of-range give INT_MIN */
int orc_trunc_int(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT_MIN;
    return (int)x;
}

/* ============================================================== window */

/* kernel/posterior_kernel.cu:635-866 (posterior_kernel + posterior_kernel2) for one
 * non-N site.  st carries the site counts and lk; rec are its sorted records;
 * next_q_sum is the following site's q_sum (read when base codes are 4..7);
 * sn is the site's known-SNP entry when refine mode applies. */
void orc_posterior_site(orc_site* st, const uint32_t* rec, int n, const int32_t* next_q_sum, const orc_params* p,
                        const double* p_prior, const orc_snp* sn, const double* p_rank)
{
    /* posterior_kernel :684-743 */
    int base1 = 0, base2 = 0;
    int qual1 = -1, qual2 = -2, qual3 = -3;
    int ac1 = 0, ac2 = 0, ac3 = 0;
    if (st->dep_uni) {
        for (int i = 0; i != 4; i++) {
            int v = st->q_sum[i];
            if (v >= qual1) {
                qual3 = qual2;
                base2 = base1;
                qual2 = qual1;
                base1 = i;
                qual1 = v;
            } else if (v >= qual2) {
                qual3 = qual2;
                base2 = i;
                qual2 = v;
            } else if (v >= qual3) {
                qual3 = v;
            }
        }
        if (qual1 == 0)
            base1 = st->ori & 7;
        else if (qual2 == 0 && base1 != (st->ori & 7))
            base2 = st->ori & 7;
    } else {
        for (int i = 0; i != 4; i++) {
            int v = st->cnt_all[i];
            if (v >= ac1) {
                ac3 = ac2;
                base2 = base1;
                ac2 = ac1;
                base1 = i;
                ac1 = v;
            } else if (v >= ac2) {
                ac3 = ac2;
                base2 = i;
                ac2 = v;
            } else if (v >= ac3) {
                ac3 = v;
            }
        }
        if (ac1 == 0)
            base1 = st->ori & 7;
        else if (ac2 == 0 && base1 != (st->ori & 7))
            base2 = st->ori & 7;
    }
    /* :746-775 */
    double prior[16];
    for (int i = 0; i < 16; i++) prior[i] = p_prior[((st->ori & 7) << 4) + i];
    if (sn) orc_snp_prior(prior, st->ori, sn, p);
    double tp[17];
    for (int i = 0; i < 17; i++) tp[i] = 0.0;
    int type1 = 16, type2 = 16;
    int gi = 0;
    for (int a1 = 0; a1 != 4; a1++) {
        for (int a2 = a1; a2 != 4; a2++, gi++) {
            int gt = a1 << 2 | a2;
            if (p->is_monoploid && a1 != a2) continue;
            tp[gt] = st->lk[gi] + log10(prior[gt]);
            if (tp[gt] >= tp[type1] || type1 == 16) {
                type2 = type1;
                type1 = gt;
            } else if (tp[gt] >= tp[type2] || type2 == 16) {
                type2 = gt;
            }
        }
    }
    /* posterior_kernel2 :820-866 */
    double rank = p->rank_sum_mode ? orc_rank_test(rec, n, st->cnt_uni, p->q_max - p->q_min, type1, p_rank) : 1.0;
    int q_cns;
    if (rank == 0.0)
        q_cns = 0;
    else
        q_cns = orc_trunc_int(10 * (tp[type1] - tp[type2]) + 10 * log10(rank));
    /* q_sum is indexed flat over the window, so base codes 4..7 (N reference)
     * read the next site's q_sum (kernel/posterior_kernel.cu:848-851) */
    if ((type1 & 3) == ((type1 >> 2) & 3)) {
        if (qual1 > 0 && base1 != (type1 & 3))
            q_cns = 0;
        else if (q_cns > qual1 - qual2)
            q_cns = qual1 - qual2;
    } else {
        int qb1 = base1 < 4 ? st->q_sum[base1] : next_q_sum[base1 - 4];
        int qb2 = base2 < 4 ? st->q_sum[base2] : next_q_sum[base2 - 4];
        int het = base1 < base2 ? (base1 << 2 | base2) : (base2 << 2 | base1);
        if (qb1 > 0 && qb2 > 0 && type1 == het) {
            if (q_cns > qual2 - qual3) q_cns = qual2 - qual3;
        } else {
            q_cns = 0;
        }
    }
    if (q_cns > 99) q_cns = 99;
    if (q_cns < 0) q_cns = 0;
    st->base1 = (uint8_t)base1;
    st->base2 = (uint8_t)base2;
    st->type1 = (uint8_t)type1;
    st->type2 = (uint8_t)type2;
    st->qual1 = qual1;
    st->qual2 = qual2;
    st->qual3 = qual3;
    st->q_cns = q_cns;
    st->rank_sum = rank;
}

/* One window: kernel/update_kernel.cu:96-195 (gold_updateSparse semantics,
 * per-read order), :233-326 (site gather); sort; likelihood; posterior
 * kernel/posterior_kernel.cu:635-866. */
int orc_call_window(const orc_reads* r, const int64_t* idx, int64_t nidx, const orc_chr* c, uint32_t ws, uint32_t C,
                    const orc_params* p, const double* S, const double* p_prior, const double* p_rank, orc_site* out,
                    uint32_t* rec_out, int32_t* nrec_out)
{
    const int L = r->L;
    const uint32_t NS = C + 1;
    memset(out, 0, sizeof(orc_site) * NS);
    uint32_t* rec = rec_out ? rec_out : (uint32_t*)malloc(sizeof(uint32_t) * (size_t)NS * ORC_MAX_REC);
    if (!rec) return fail(ORC_E_NOMEM, "records");
    /* idx lists the chromosome's reads that pass pos >= 0 and pos+len < chr_len
     * (cpp/call_genotype.cc:495-514), in input order, positions non-decreasing.
     * First read that can cover ws: pos >= ws - L + 1. */
    int64_t lo = 0, hi = nidx;
    int64_t want = (int64_t)ws - L + 1;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (r->pos[idx[mid]] < want)
            lo = mid + 1;
        else
            hi = mid;
    }
    for (int64_t k = lo; k < nidx && r->pos[idx[k]] <= (int64_t)ws + C; k++) {
        int64_t i = idx[k];
        const char* s = r->seq + (size_t)i * L;
        const char* q = r->qual + (size_t)i * L;
        for (int coord = 0; coord < r->len[i]; coord++) {
            int64_t site = (int64_t)r->pos[i] + coord - ws;
            if (site < 0 || site >= NS) continue;
            /* the extra site C of a full window is the first tail site, which at
             * call time holds only reads that start inside the window
             * (kernel/update_kernel.cu:269-274 stops at global_win_size) */
            if (site == C && r->pos[i] >= (int64_t)ws + ORC_WIN) continue;
            orc_site* st = &out[site];
            st->depth += 1;
            st->repeat += r->hit[i];
            int qc = (unsigned char)q[coord];
            if (!is_acgt(s[coord]) || qc < p->q_min || st->dep_uni >= 0xFF) continue;
            int o = ((unsigned)s[coord] >> 1) & 3;
            if (r->hit[i] == 1) {
                int qq = qc - p->q_min;
                uint32_t strand = r->strand[i] == '+' ? 0 : 1;
                uint32_t cyc = strand ? (uint32_t)(r->len[i] - 1 - coord) : (uint32_t)coord;
                rec[(size_t)site * ORC_MAX_REC + st->dep_uni] = orc_encode_word(cyc, strand, o, qq, 1);
                st->dep_uni += 1;
                st->cnt_uni[o] += 1;
                st->q_sum[o] += qq;
            }
            st->cnt_all[o] += 1;
        }
    }
    for (uint32_t j = 0; j < C; j++) {
        orc_site* st = &out[j];
        uint32_t* rr = rec + (size_t)j * ORC_MAX_REC;
        int n = st->dep_uni;
        orc_sort_records(rr, n);
        if (nrec_out) nrec_out[j] = n;
        st->ori = c->code[ws + j] & 0xF; /* cpp/call_genotype.cc:222-230 */
        st->nflag = ((st->ori & 4) != 0) && st->depth == 0;
        int rc = orc_likelihood_sorted(rr, n, L, p->pcr_dependency, p->global_dependency, S, st->lk);
        if (rc) {
            if (!rec_out) free(rec);
            return rc;
        }
        if (st->nflag) continue;
        const orc_snp* sn = ((st->ori & 0x8) && p->refine_mode) ? orc_find_snp(c, ws + j) : NULL;
        orc_posterior_site(st, rr, n, out[j + 1].q_sum, p, p_prior, sn, p_rank);
    }
    if (!rec_out) free(rec);
    return ORC_OK;
}

/* =============================================================== text */

int orc_format_g(double v, char* out) { return sprintf(out, "%#g", v); }

static const char ABBV[17] = {'A', 'M', 'W', 'R', 'M', 'C', 'Y', 'S', 'W', 'Y', 'T', 'K', 'R', 'S', 'K', 'G', 'N'};
static const char ACTGN[9] = "ACTGNNNN";

/* cpp/call_genotype.cc:353-411 (ostream rows; showpoint doubles are %#g) */
int orc_format_row(const char* chr, uint32_t pos1, const orc_site* s, const orc_site* next, char* o)
{
    (void)next;
    if (s->nflag || s->base1 >= 4)
        return sprintf(o, "%s\t%u\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n", chr, pos1);
    int b1 = s->base1, b2 = s->base2;
    int avg1 = s->q_sum[b1] == 0 ? 0 : s->q_sum[b1] / s->cnt_uni[b1];
    char rs[64], cp[64];
    sprintf(rs, "%#g", s->rank_sum);
    sprintf(cp, "%#g", s->depth == 0 ? 255.0 : (double)s->repeat / s->depth);
    int n = sprintf(o, "%s\t%u\t%c\t%c\t%d\t%c\t%d\t%d\t%d\t", chr, pos1, ACTGN[s->ori & 7], ABBV[s->type1], s->q_cns,
                    ACTGN[b1], avg1, s->cnt_uni[b1], s->cnt_all[b1]);
    if (b2 < 4) {
        int avg2 = s->q_sum[b2] == 0 ? 0 : s->q_sum[b2] / s->cnt_uni[b2];
        n += sprintf(o + n, "%c\t%d\t%d\t%d\t", ACTGN[b2], avg2, s->cnt_uni[b2], s->cnt_all[b2]);
    } else {
        n += sprintf(o + n, "N\t0\t0\t0\t");
    }
    n += sprintf(o + n, "%d\t%s\t%s\t%d\n", s->depth, rs, cp, (s->ori & 8) ? 1 : 0);
    return n;
}

/* =========================================================== -C codec */

/* cpp/dict_compression.h:47-54 */
int orc_num_bits(uint64_t n)
{
    int b = (int)ceil(log((double)(n)) / log((double)2.f));
    if (b == 0) b = 1;
    return b;
}

/* cpp/dict_compression.h:116-166 dictCompression (sorted unique dictionary, MSB-first
 * packing, last word holds the remainder right-aligned).  n==0 writes nothing. */
#define DICT_ENCODE_BODY(T, cmp)                                                     \
    if (n == 0) {                                                                    \
        *n_dict = 0;                                                                 \
        return 0;                                                                    \
    }                                                                                \
    memcpy(dict, in, sizeof(T) * n);                                                 \
    qsort(dict, n, sizeof(T), cmp);                                                  \
    uint64_t k = 1;                                                                  \
    for (uint64_t i = 1; i < n; i++)                                                 \
        if (!(dict[i] == dict[k - 1])) dict[k++] = dict[i];                          \
    *n_dict = k;                                                                     \
    int bits = orc_num_bits(k);                                                      \
    int per = (int)floor(32 / (double)bits);                                         \
    uint64_t nw = (uint64_t)ceil(n / (double)per);                                   \
    memset(words, 0, sizeof(uint32_t) * nw);                                         \
    for (uint64_t i = 0; i < n; i++) {                                               \
        uint64_t lo = 0, hi = k;                                                     \
        while (lo < hi) {                                                            \
            uint64_t mid = (lo + hi) / 2;                                            \
            if (dict[mid] < in[i])                                                   \
                lo = mid + 1;                                                        \
            else                                                                     \
                hi = mid;                                                            \
        }                                                                            \
        uint64_t w = i / per;                                                        \
        words[w] = bits == 32 ? 0 : (words[w] << bits);                              \
        words[w] |= (uint32_t)lo;                                                    \
    }                                                                                \
    return nw;

static int i32_cmp(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return x < y ? -1 : (x > y);
}
static int f64_cmp(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y);
}
uint64_t orc_dict_encode_i32(const int32_t* in, uint64_t n, uint32_t* words, int32_t* dict, uint64_t* n_dict)
{
    DICT_ENCODE_BODY(int32_t, i32_cmp)
}
uint64_t orc_dict_encode_u32(const uint32_t* in, uint64_t n, uint32_t* words, uint32_t* dict, uint64_t* n_dict)
{
    DICT_ENCODE_BODY(uint32_t, u32_cmp)
}
uint64_t orc_dict_encode_f64(const double* in, uint64_t n, uint32_t* words, double* dict, uint64_t* n_dict)
{
    DICT_ENCODE_BODY(double, f64_cmp)
}

/* kernel/compress_kernel.cu:104-153: A0 C1 G2 T3, 16 per word, MSB first */
static uint32_t base2int(char b)
{
    switch (b) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}
uint64_t orc_encode_2bit(const char* seq, uint64_t n, uint32_t* code)
{
    uint64_t nw = (uint64_t)ceil(n / 16.0); /* getNumBinWord :85-90 */
    for (uint64_t w = 0; w < nw; w++) {
        uint64_t m = (w == nw - 1) ? n - w * 16 : 16;
        uint32_t x = 0;
        for (uint64_t b = 0; b < m; b++) x = (x << 2) | base2int(seq[w * 16 + b]);
        code[w] = x;
    }
    return nw;
}
void orc_decode_2bit(char* seq, uint64_t n, const uint32_t* code, uint64_t nw)
{
    static const char T[4] = {'A', 'C', 'G', 'T'};
    if (nw == 0) return;
    for (uint64_t w = 0; w + 1 < nw; w++)
        for (int b = 0; b < 16; b++) seq[w * 16 + b] = T[(code[w] >> (30 - 2 * b)) & 3];
    uint64_t left = n - (nw - 1) * 16;
    for (uint64_t b = 0; b < left; b++) seq[(nw - 1) * 16 + b] = T[(code[nw - 1] >> (2 * (left - b - 1))) & 3];
}

/* ------------------------------------------------------------ -C writer */
typedef struct {
    FILE *f346, *f5, *f10, *f11, *f15;
    /* .1_2 index */
    char** names;
    uint64_t* starts;
    uint64_t* sizes;
    uint64_t nwin, capwin;
} cwriter;

static void wr(FILE* f, const void* p, size_t sz, size_t n)
{
    if (n) fwrite(p, sz, n, f);
}

static int all_n(const orc_site* s, uint32_t C)
{
    for (uint32_t j = 0; j < C; j++)
        if (!s[j].nflag) return 0;
    return 1;
}

/* RLE (kernel/compress_kernel.cu:336-393) + dictionaries (cpp/dict_compression.h:235-270),
 * one column as written by Column_RLEDict::compress_write (cpp/compress.h:501-610) */
static void write_rle_i32(FILE* f, const int32_t* v, uint32_t C)
{
    int32_t* rv = (int32_t*)malloc(sizeof(int32_t) * C);
    uint32_t* rl = (uint32_t*)malloc(sizeof(uint32_t) * C);
    uint64_t nr = 0;
    for (uint32_t i = 0; i < C; i++) {
        if (nr && rv[nr - 1] == v[i])
            rl[nr - 1]++;
        else {
            rv[nr] = v[i];
            rl[nr] = 1;
            nr++;
        }
    }
    uint32_t* wv = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    uint32_t* wl = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    int32_t* dv = (int32_t*)malloc(sizeof(int32_t) * (nr + 1));
    uint32_t* dl = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    uint64_t ndv, ndl;
    uint64_t nwv = orc_dict_encode_i32(rv, nr, wv, dv, &ndv);
    uint64_t nwl = orc_dict_encode_u32(rl, nr, wl, dl, &ndl);
    uint64_t param[6] = {nr, nwv, nwl, ndv, ndl, 4 * nwv + 4 * nwl + 4 * ndv + 4 * ndl};
    wr(f, param, 8, 6);
    wr(f, wv, 4, nwv);
    wr(f, wl, 4, nwl);
    wr(f, dv, 4, ndv);
    wr(f, dl, 4, ndl);
    free(rv), free(rl), free(wv), free(wl), free(dv), free(dl);
}

static void write_rle_f64(FILE* f, const double* v, uint32_t C)
{
    double* rv = (double*)malloc(sizeof(double) * C);
    uint32_t* rl = (uint32_t*)malloc(sizeof(uint32_t) * C);
    uint64_t nr = 0;
    for (uint32_t i = 0; i < C; i++) {
        if (nr && rv[nr - 1] == v[i])
            rl[nr - 1]++;
        else {
            rv[nr] = v[i];
            rl[nr] = 1;
            nr++;
        }
    }
    uint32_t* wv = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    uint32_t* wl = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    double* dv = (double*)malloc(sizeof(double) * (nr + 1));
    uint32_t* dl = (uint32_t*)malloc(sizeof(uint32_t) * (nr + 1));
    uint64_t ndv, ndl;
    uint64_t nwv = orc_dict_encode_f64(rv, nr, wv, dv, &ndv);
    uint64_t nwl = orc_dict_encode_u32(rl, nr, wl, dl, &ndl);
    uint64_t param[6] = {nr, nwv, nwl, ndv, ndl, 4 * nwv + 4 * nwl + 8 * ndv + 4 * ndl};
    wr(f, param, 8, 6);
    wr(f, wv, 4, nwv);
    wr(f, wl, 4, nwl);
    wr(f, dv, 8, ndv);
    wr(f, dl, 4, ndl);
    free(rv), free(rl), free(wv), free(wl), free(dv), free(dl);
}

/* one window chunk of each column file (cpp/call_genotype.cc:312-352, cpp/compress.h) */
static void cwrite_window(cwriter* w, const char* chr, uint32_t ws, uint32_t C, const orc_site* s)
{
    if (w->nwin == w->capwin) {
        w->capwin = w->capwin ? w->capwin * 2 : 64;
        w->names = (char**)realloc(w->names, sizeof(char*) * w->capwin);
        w->starts = (uint64_t*)realloc(w->starts, sizeof(uint64_t) * w->capwin);
        w->sizes = (uint64_t*)realloc(w->sizes, sizeof(uint64_t) * w->capwin);
    }
    w->names[w->nwin] = strdup(chr);
    w->starts[w->nwin] = (uint64_t)ws + 1;
    w->sizes[w->nwin] = C;
    w->nwin++;
    int allN = all_n(s, C);
    /* .3_4_6 (cpp/compress.h:187-409) */
    if (allN) {
        uint64_t param[7] = {1, 0, C, 0, 0, 0, 0};
        wr(w->f346, param, 8, 7);
    } else {
        char *c3 = (char*)malloc(C + 1), *c4 = (char*)malloc(C + 1), *c6 = (char*)malloc(C + 1);
        for (uint32_t j = 0; j < C; j++) {
            if (s[j].nflag || s[j].base1 >= 4) {
                c3[j] = c4[j] = c6[j] = 'N';
            } else {
                c3[j] = ACTGN[s[j].ori & 7];
                c4[j] = ABBV[s[j].type1];
                c6[j] = ACTGN[s[j].base1];
            }
        }
        uint64_t n3 = 0, n4 = 0, n6 = 0;
        for (uint32_t j = 0; j < C; j++) {
            n3 += c3[j] == 'N';
            n4 += c4[j] != c3[j];
            n6 += c6[j] != c3[j];
        }
        uint64_t ncode = (uint64_t)ceil(C / 16.0);
        uint64_t size = 4 * n4 + n4 + 4 * n6 + n6 + 4 * n3 + 4 * ncode;
        uint64_t param[7] = {0, size, C, ncode, n3, n4, n6};
        wr(w->f346, param, 8, 7);
        int32_t* id4 = (int32_t*)malloc(4 * (n4 + 1));
        char* v4 = (char*)malloc(n4 + 1);
        int32_t* id6 = (int32_t*)malloc(4 * (n6 + 1));
        char* v6 = (char*)malloc(n6 + 1);
        int32_t* idn = (int32_t*)malloc(4 * (n3 + 1));
        uint64_t k4 = 0, k6 = 0, k3 = 0;
        for (uint32_t j = 0; j < C; j++) {
            if (c4[j] != c3[j]) id4[k4] = j, v4[k4++] = c4[j];
            if (c6[j] != c3[j]) id6[k6] = j, v6[k6++] = c6[j];
            if (c3[j] == 'N') idn[k3++] = j, c3[j] = 'A';
        }
        uint32_t* code = (uint32_t*)malloc(4 * (ncode + 1));
        orc_encode_2bit(c3, C, code);
        wr(w->f346, id4, 4, n4);
        wr(w->f346, v4, 1, n4);
        wr(w->f346, id6, 4, n6);
        wr(w->f346, v6, 1, n6);
        wr(w->f346, idn, 4, n3);
        wr(w->f346, code, 4, ncode);
        free(c3), free(c4), free(c6), free(id4), free(v4), free(id6), free(v6), free(idn), free(code);
    }
    int32_t nind = allN ? 1 : 0;
    /* .5_7_8_9_14_16 */
    wr(w->f5, &nind, 4, 1);
    if (!allN) {
        int32_t* col = (int32_t*)malloc(4 * (C + 1));
        double* cold = (double*)malloc(8 * (C + 1));
        for (int which = 0; which < 5; which++) {
            for (uint32_t j = 0; j < C; j++) {
                const orc_site* t = &s[j];
                int32_t v = 0;
                if (!t->nflag && t->base1 < 4) {
                    int b = t->base1;
                    switch (which) {
                    case 0: v = t->q_cns; break;
                    case 1: v = t->q_sum[b] == 0 ? 0 : t->q_sum[b] / t->cnt_uni[b]; break;
                    case 2: v = t->cnt_uni[b]; break;
                    case 3: v = t->cnt_all[b]; break;
                    default: v = t->depth; break;
                    }
                }
                col[j] = v;
            }
            write_rle_i32(w->f5, col, C);
        }
        for (uint32_t j = 0; j < C; j++) {
            const orc_site* t = &s[j];
            cold[j] = (!t->nflag && t->base1 < 4) ? (t->depth == 0 ? 255 : (double)t->repeat / t->depth) : 255.0;
        }
        write_rle_f64(w->f5, cold, C);
        free(col);
        free(cold);
    }
    /* .10 (cpp/compress.h:786-897) */
    wr(w->f10, &nind, 4, 1);
    if (!allN) {
        char* c10 = (char*)malloc(C + 1);
        uint32_t* nidx = (uint32_t*)malloc(4 * (C + 1));
        uint64_t nn = 0;
        for (uint32_t j = 0; j < C; j++) {
            c10[j] = (!s[j].nflag && s[j].base1 < 4 && s[j].base2 < 4) ? ACTGN[s[j].base2] : 'N';
            if (c10[j] == 'N') nidx[nn++] = j, c10[j] = 'A';
        }
        uint64_t ncode = (uint64_t)ceil(C / 16.0);
        uint32_t* code = (uint32_t*)malloc(4 * (ncode + 1));
        orc_encode_2bit(c10, C, code);
        uint64_t param[3] = {nn, ncode, C};
        wr(w->f10, param, 8, 3);
        wr(w->f10, code, 4, ncode);
        wr(w->f10, nidx, 4, nn);
        free(c10), free(nidx), free(code);
    }
    /* .11_12_13 (cpp/compress.h:951-1175); the else-if chain drops col 13 when
     * col11==0, col12!=0 (reproduced) */
    wr(w->f11, &nind, 4, 1);
    if (!allN) {
        int32_t *cid = (int32_t*)malloc(4 * (C + 1)), *c11 = (int32_t*)malloc(4 * (C + 1)),
                *c12 = (int32_t*)malloc(4 * (C + 1)), *c13 = (int32_t*)malloc(4 * (C + 1));
        int32_t *x12i = (int32_t*)malloc(4 * (C + 1)), *x12v = (int32_t*)malloc(4 * (C + 1)),
                *x13i = (int32_t*)malloc(4 * (C + 1)), *x13v = (int32_t*)malloc(4 * (C + 1));
        uint64_t nc = 0, n12 = 0, n13 = 0;
        for (uint32_t j = 0; j < C; j++) {
            const orc_site* t = &s[j];
            int32_t a = 0, b = 0, d = 0;
            if (!t->nflag && t->base1 < 4 && t->base2 < 4) {
                int b2 = t->base2;
                a = t->q_sum[b2] == 0 ? 0 : t->q_sum[b2] / t->cnt_uni[b2];
                b = t->cnt_uni[b2];
                d = t->cnt_all[b2];
            }
            if (a != 0) {
                cid[nc] = j, c11[nc] = a, c12[nc] = b, c13[nc] = d, nc++;
            } else if (a != b) {
                x12i[n12] = j, x12v[n12++] = b;
            } else if (a != d) {
                x13i[n13] = j, x13v[n13++] = d;
            }
        }
        uint64_t param[4] = {C, nc, n12, n13};
        wr(w->f11, param, 8, 4);
        wr(w->f11, x12i, 4, n12);
        wr(w->f11, x12v, 4, n12);
        wr(w->f11, x13i, 4, n13);
        wr(w->f11, x13v, 4, n13);
        wr(w->f11, cid, 4, nc);
        uint32_t* words = (uint32_t*)malloc(4 * (nc + 1));
        int32_t* dict = (int32_t*)malloc(4 * (nc + 1));
        int32_t* cols[3] = {c11, c12, c13};
        for (int k = 0; k < 3; k++) {
            uint64_t nd;
            uint64_t nw = orc_dict_encode_i32(cols[k], nc, words, dict, &nd);
            uint64_t dp[2] = {nw, nd};
            wr(w->f11, dp, 8, 2);
            wr(w->f11, words, 4, nw);
            wr(w->f11, dict, 4, nd);
        }
        free(cid), free(c11), free(c12), free(c13), free(x12i), free(x12v), free(x13i), free(x13v);
        free(words), free(dict);
    }
    /* .15_17 (cpp/compress.h:1178-1312) */
    wr(w->f15, &nind, 4, 1);
    if (!allN) {
        int32_t* id = (int32_t*)malloc(4 * (C + 1));
        double* val = (double*)malloc(8 * (C + 1));
        int32_t n = 0;
        for (uint32_t j = 0; j < C; j++)
            if (!s[j].nflag && s[j].base1 < 4 && s[j].rank_sum != 1.0) id[n] = j, val[n++] = s[j].rank_sum;
        wr(w->f15, &n, 4, 1);
        wr(w->f15, id, 4, n);
        wr(w->f15, val, 8, n);
        n = 0;
        for (uint32_t j = 0; j < C; j++)
            if (!s[j].nflag && s[j].base1 < 4 && (s[j].ori & 8)) id[n++] = j;
        wr(w->f15, &n, 4, 1);
        wr(w->f15, id, 4, n);
        free(id);
        free(val);
    }
}

/* ============================================================ regions */

/* Chr_info::set_region, cpp/chromosome.cc:66-107: clip [start, end] to
 * [0, len], then set sites start..end (inclusive) in the site mask and
 * windows start/W..end/W in the window mask; words of 64 positions, position
 * p at bit 63 - p % 64.  (The reference's shift by 64 when a range fills one
 * whole word sets nothing on x86; the intended bits are set here.) */
static void mask_set(uint64_t* m, int64_t a, int64_t b)
{
    for (int64_t p = a; p <= b; p++) m[p / 64] |= 1ull << (63 - p % 64);
}

int orc_region_set(uint64_t* site_mask, uint64_t* win_mask, uint32_t len, int start, int end)
{
    if (start < 0)
        start = 0;
    else if ((uint32_t)start >= len)
        start = (int)len;
    if (end < 0)
        end = 0;
    else if ((uint32_t)end >= len)
        end = (int)len;
    if (start > end) return fail(ORC_E_FORMAT, "Invalid region: %d-%d", start, end);
    if (site_mask) mask_set(site_mask, start, end);
    if (win_mask) mask_set(win_mask, start / ORC_WIN, end / ORC_WIN);
    return ORC_OK;
}

/* Chr_info::is_in_region_win, cpp/soap_snp.h:796-800 */
int orc_region_in_win(const uint64_t* win_mask, uint32_t pos)
{
    const uint32_t w = pos / ORC_WIN;
    return (int)((win_mask[w / 64] >> (63 - w % 64)) & 1);
}

/* Words of a mask over n positions (one spare position: set_region clips the
 * end to len, not len - 1). */
uint64_t orc_region_words(uint64_t n) { return n / 64 + 1; }

/* Genome::read_region, cpp/chromosome.cc:127-170: lines "chr start end"
 * (1-based, inclusive); each sets set_region(start - read_length, end - 1) on
 * its chromosome, whose masks start empty; lines of unknown chromosomes are
 * skipped.  win_masks[c] stays NULL for a chromosome without a line. */
static int read_regions(const char* path, const orc_genome* g, int read_length, uint64_t** win_masks)
{
    FILE* f = fopen(path, "r");
    if (!f) return fail(ORC_E_IO, "cannot open %s", path);
    char line[4096], name[1024];
    long long a, b;
    int rc = ORC_OK;
    while (!rc && fgets(line, sizeof line, f)) {
        if (sscanf(line, "%1023s %lld %lld", name, &a, &b) != 3) {
            rc = fail(ORC_E_FORMAT, "Wrong format in target region file");
            break;
        }
        int ci = orc_genome_find(g, name);
        if (ci < 0) {
            fprintf(stderr, "Unexpected Chromosome:%s\n", name);
            continue;
        }
        const orc_chr* c = &g->chr[ci];
        if (!win_masks[ci])
            win_masks[ci] = (uint64_t*)calloc(orc_region_words(c->len / ORC_WIN + 1), sizeof(uint64_t));
        if (a - read_length < INT_MIN || b - 1 > INT_MAX) {
            rc = fail(ORC_E_FORMAT, "region bounds out of range");
            break;
        }
        rc = orc_region_set(NULL, win_masks[ci], c->len, (int)(a - read_length), (int)(b - 1));
    }
    fclose(f);
    return rc;
}

/* ============================================================ pipeline */

static FILE* open_sfx(const char* prefix, const char* sfx, const char* mode);

/* .1_2 (cpp/compress.h:151-165): window count, then "chr\tposStart\twinSize" per window */
static int cwrite_index(const cwriter* w, const char* prefix)
{
    FILE* f12 = open_sfx(prefix, ".1_2", "w");
    if (!f12) return fail(ORC_E_IO, "cannot create .1_2");
    fprintf(f12, "%lu\n", (unsigned long)w->nwin);
    for (uint64_t i = 0; i < w->nwin; i++)
        fprintf(f12, "%s\t%lu\t%lu\n", w->names[i], (unsigned long)w->starts[i], (unsigned long)w->sizes[i]);
    fclose(f12);
    return ORC_OK;
}

static int cwriter_open(cwriter* w, const char* prefix)
{
    memset(w, 0, sizeof *w);
    w->f346 = open_sfx(prefix, ".3_4_6", "wb");
    w->f5 = open_sfx(prefix, ".5_7_8_9_14_16", "wb");
    w->f10 = open_sfx(prefix, ".10", "wb");
    w->f11 = open_sfx(prefix, ".11_12_13", "wb");
    w->f15 = open_sfx(prefix, ".15_17", "wb");
    if (!w->f346 || !w->f5 || !w->f10 || !w->f11 || !w->f15) return fail(ORC_E_IO, "cannot create -C files");
    return ORC_OK;
}

static void cwriter_close(cwriter* w)
{
    if (w->f346) fclose(w->f346);
    if (w->f5) fclose(w->f5);
    if (w->f10) fclose(w->f10);
    if (w->f11) fclose(w->f11);
    if (w->f15) fclose(w->f15);
    for (uint64_t i = 0; i < w->nwin; i++) free(w->names[i]);
    free(w->names), free(w->starts), free(w->sizes);
    memset(w, 0, sizeof *w);
}

/* The -C files of one window of given sites (the chunk writer on its own,
 * for byte-level known-answer tests of the layout). */
int orc_cwrite_sites(const char* prefix, const char* chr, uint32_t ws, uint32_t C, const orc_site* sites)
{
    cwriter cw;
    int rc = cwriter_open(&cw, prefix);
    if (!rc) {
        cwrite_window(&cw, chr, ws, C, sites);
        rc = cwrite_index(&cw, prefix);
    }
    cwriter_close(&cw);
    return rc;
}



static FILE* open_sfx(const char* prefix, const char* sfx, const char* mode)
{
    char path[4096];
    snprintf(path, sizeof path, "%s%s", prefix, sfx);
    return fopen(path, mode);
}

static double now_sec(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

/* cpp/main.cc:91-429 + cpp/call_genotype.cc:422-669, extended to several
 * chromosomes called in order of first appearance (SOAPsnp style); the
 * window list of a chromosome of length N is N/W+1 windows, the last one
 * of N%W sites (cpp/call_genotype.cc:592-644).  max_sites > 0 stops after
 * that many called sites (a bounded sample for timing). */
/* Text of one window's rows, built by a worker of the threaded run. */
typedef struct {
    char* text;
    size_t len;
    int rc, ready;
} win_text;

typedef struct {
    const orc_reads* r;
    const orc_genome* g;
    const orc_params* p;
    const double *S, *pp, *pr;
    const int64_t* idx;
    const int *wblk;     /* per window: chromosome block */
    const int64_t *b0, *b1;
    const int* order;
    const uint32_t *wstart, *wsize;
    int nw;
    int next;            /* next window to take */
    int written;         /* windows handed to the writer so far */
    int ahead;           /* at most this many windows past `written` in flight */
    win_text* out;
    pthread_mutex_t mu;
    pthread_cond_t cv;
} run_pool;

static void* run_worker(void* arg)
{
    run_pool* P = (run_pool*)arg;
    orc_site* sites = (orc_site*)malloc(sizeof(orc_site) * (ORC_WIN + 1));
    uint32_t* rec = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ORC_WIN + 1) * ORC_MAX_REC);
    char row[4096];
    for (;;) {
        pthread_mutex_lock(&P->mu);
        while (P->next < P->nw && P->next >= P->written + P->ahead) pthread_cond_wait(&P->cv, &P->mu);
        int w = P->next < P->nw ? P->next++ : -1;
        pthread_mutex_unlock(&P->mu);
        if (w < 0) break;
        const int b = P->wblk[w];
        const orc_chr* c = &P->g->chr[P->order[b]];
        const uint32_t ws = P->wstart[w], C = P->wsize[w];
        win_text t = {NULL, 0, ORC_OK, 1};
        if (!sites || !rec) {
            t.rc = ORC_E_NOMEM;
        } else {
            t.rc = orc_call_window(P->r, P->idx + P->b0[b], P->b1[b] - P->b0[b], c, ws, C, P->p, P->S, P->pp, P->pr,
                                   sites, rec, NULL);
        }
        if (!t.rc) {
            size_t cap = (size_t)C * 96 + 64, len = 0;
            t.text = (char*)malloc(cap);
            for (uint32_t j = 0; t.text && j < C; j++) {
                int n = orc_format_row(c->name, ws + j + 1, &sites[j], &sites[j + 1], row);
                if (len + (size_t)n > cap) {
                    cap = cap * 2 + (size_t)n;
                    t.text = (char*)realloc(t.text, cap);
                    if (!t.text) break;
                }
                memcpy(t.text + len, row, (size_t)n);
                len += (size_t)n;
            }
            if (!t.text) t.rc = ORC_E_NOMEM;
            t.len = len;
        }
        pthread_mutex_lock(&P->mu);
        P->out[w] = t;
        pthread_cond_broadcast(&P->cv);
        pthread_mutex_unlock(&P->mu);
    }
    free(sites);
    free(rec);
    return NULL;
}

int orc_run(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
            const orc_params* p, const char* out, int compress, int64_t max_sites, orc_run_stats* stats)
{
    return orc_run_opts(fasta, soap, dbsnp, matrix_in, matrix_out, p, out, compress, max_sites, 1, NULL, stats);
}

int orc_run_threads(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
                    const char* matrix_out, const orc_params* p, const char* out, int compress, int64_t max_sites,
                    int nthreads, orc_run_stats* stats)
{
    return orc_run_opts(fasta, soap, dbsnp, matrix_in, matrix_out, p, out, compress, max_sites, nthreads, NULL,
                        stats);
}

/* orc_run with the windows of the text output called on nthreads threads
 * (each window is independent given the reads: orc_call_window builds its
 * site matrix from every read that covers it), rows written in window order. */
/* With regions (-T), only the windows that a region touches are called and
 * written (the window gating of cpp/call_genotype.cc:535 and :593); each is
 * built from every read that covers it. */
int orc_run_opts(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
                 const char* matrix_out, const orc_params* p, const char* out, int compress, int64_t max_sites,
                 int nthreads, const char* regions, orc_run_stats* stats)
{
    orc_genome* g = NULL;
    orc_reads* r = NULL;
    int rc = orc_genome_load(fasta, dbsnp, &g);
    if (rc) return rc;
    rc = orc_reads_load(soap, p, g, &r);
    if (rc) {
        orc_genome_free(g);
        return rc;
    }
    double* pm = (double*)malloc(sizeof(double) * ORC_PMATRIX);
    double* S = (double*)malloc(sizeof(double) * ORC_SCORE);
    double* pp = (double*)malloc(sizeof(double) * 128);
    double* pr = (double*)malloc(sizeof(double) * ORC_PRANK);
    if (matrix_in && matrix_in[0])
        rc = orc_matrix_read(matrix_in, pm);
    else
        rc = orc_matrix_gen(r, g, p, pm);
    if (!rc && matrix_out && matrix_out[0]) rc = orc_matrix_write(matrix_out, pm, p);
    /* -T: window masks per chromosome (NULL: no region line, nothing called) */
    uint64_t** wmask = NULL;
    if (!rc && regions && regions[0]) {
        wmask = (uint64_t**)calloc((size_t)g->n + 1, sizeof(uint64_t*));
        rc = read_regions(regions, g, p->read_length, wmask);
    }
#define ORC_CALLED(ci, ws) (!wmask || (wmask[ci] && orc_region_in_win(wmask[ci], (ws))))
    orc_prior_gen(p, pp);
    orc_rank_table_gen(pr);
    orc_score_table(pm, S);
    /* chromosome blocks in order of first appearance (reads with pos >= 0);
     * each block keeps the reads that pass pos+len < chr_len, whose positions
     * must not decrease (cpp/call_genotype.cc:495-519) */
    int* order = (int*)malloc(sizeof(int) * (g->n + 1));
    int64_t* b0 = (int64_t*)malloc(sizeof(int64_t) * (g->n + 1));
    int64_t* b1 = (int64_t*)malloc(sizeof(int64_t) * (g->n + 1));
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (r->n + 1));
    int64_t nidx = 0;
    int* seen = (int*)calloc(g->n, sizeof(int));
    int nblk = 0;
    int cur = -1;
    int64_t last = 0;
    for (int64_t i = 0; !rc && i < r->n; i++) {
        if (r->pos[i] < 0) continue;
        int ci = r->chr[i];
        if (ci != cur) {
            if (seen[ci]) {
                rc = fail(ORC_E_UNSORTED, "reads of chromosome %s are not contiguous", g->chr[ci].name);
                break;
            }
            seen[ci] = 1;
            if (nblk) b1[nblk - 1] = nidx;
            order[nblk] = ci;
            b0[nblk] = nidx;
            nblk++;
            cur = ci;
            last = 0;
        }
        if ((int64_t)r->pos[i] + r->len[i] >= (int64_t)g->chr[ci].len) continue;
        if (r->pos[i] < last) {
            rc = fail(ORC_E_UNSORTED, "Errors in sorting:%d<%lld", r->pos[i], (long long)last);
            break;
        }
        last = r->pos[i];
        idx[nidx++] = i;
    }
    if (nblk) b1[nblk - 1] = nidx;
    FILE* fo = NULL;
    cwriter cw;
    memset(&cw, 0, sizeof cw);
    if (!rc) {
        fo = fopen(out, "w");
        if (!fo) rc = fail(ORC_E_IO, "cannot create %s", out);
    }
    if (!rc && compress) rc = cwriter_open(&cw, out);
    orc_site* sites = (orc_site*)malloc(sizeof(orc_site) * (ORC_WIN + 1));
    char* row = (char*)malloc(4096);
    uint64_t nsites = 0, nwin = 0;
    double tcall = 0;
    if (!rc && nthreads > 1 && !compress) {
        /* the window list, bounded by max_sites as in the serial loop */
        int cap = 16, nw = 0;
        int* wblk = (int*)malloc(sizeof(int) * cap);
        uint32_t* wstart = (uint32_t*)malloc(sizeof(uint32_t) * cap);
        uint32_t* wsize = (uint32_t*)malloc(sizeof(uint32_t) * cap);
        uint64_t planned = 0;
        for (int b = 0; b < nblk; b++) {
            const orc_chr* c = &g->chr[order[b]];
            uint32_t n = c->len / ORC_WIN + 1;
            for (uint32_t w = 0; w < n; w++) {
                if (max_sites > 0 && (int64_t)planned >= max_sites) break;
                if (!ORC_CALLED(order[b], w * ORC_WIN)) continue;
                if (nw == cap) {
                    cap *= 2;
                    wblk = (int*)realloc(wblk, sizeof(int) * cap);
                    wstart = (uint32_t*)realloc(wstart, sizeof(uint32_t) * cap);
                    wsize = (uint32_t*)realloc(wsize, sizeof(uint32_t) * cap);
                }
                wblk[nw] = b;
                wstart[nw] = w * ORC_WIN;
                wsize[nw] = (w < c->len / ORC_WIN) ? ORC_WIN : c->len % ORC_WIN;
                planned += wsize[nw];
                nw++;
            }
        }
        run_pool P;
        memset(&P, 0, sizeof P);
        P.r = r, P.g = g, P.p = p, P.S = S, P.pp = pp, P.pr = pr, P.idx = idx, P.wblk = wblk, P.b0 = b0, P.b1 = b1;
        P.order = order, P.wstart = wstart, P.wsize = wsize, P.nw = nw, P.ahead = 2 * nthreads;
        P.out = (win_text*)calloc((size_t)nw + 1, sizeof(win_text));
        pthread_mutex_init(&P.mu, NULL);
        pthread_cond_init(&P.cv, NULL);
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
        double t0 = now_sec();
        int started = 0;
        for (int t = 0; t < nthreads; t++)
            if (pthread_create(&th[started], NULL, run_worker, &P) == 0) started++;
        if (!started) run_worker(&P);
        for (int w = 0; w < nw; w++) {
            pthread_mutex_lock(&P.mu);
            while (!P.out[w].ready) pthread_cond_wait(&P.cv, &P.mu);
            win_text t = P.out[w];
            pthread_mutex_unlock(&P.mu);
            if (!rc && t.rc) rc = fail(t.rc, "window %d failed", w);
            if (!rc && t.len && fwrite(t.text, 1, t.len, fo) != t.len) rc = fail(ORC_E_IO, "cannot write %s", out);
            free(t.text);
            P.out[w].text = NULL;
            if (!rc) {
                nsites += wsize[w];
                nwin++;
            }
            pthread_mutex_lock(&P.mu);
            P.written = w + 1;
            pthread_cond_broadcast(&P.cv);
            pthread_mutex_unlock(&P.mu);
        }
        for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
        tcall = now_sec() - t0;
        pthread_mutex_destroy(&P.mu);
        pthread_cond_destroy(&P.cv);
        free(th), free(P.out), free(wblk), free(wstart), free(wsize);
        nblk = 0; /* done: skip the serial loop */
    }
    for (int b = 0; !rc && b < nblk; b++) {
        const orc_chr* c = &g->chr[order[b]];
        uint32_t nw = c->len / ORC_WIN + 1;
        for (uint32_t w = 0; !rc && w < nw; w++) {
            if (max_sites > 0 && (int64_t)nsites >= max_sites) break;
            uint32_t ws = w * ORC_WIN;
            uint32_t C = (w < c->len / ORC_WIN) ? ORC_WIN : c->len % ORC_WIN;
            if (!ORC_CALLED(order[b], ws)) continue;
            double t0 = now_sec();
            rc = orc_call_window(r, idx + b0[b], b1[b] - b0[b], c, ws, C, p, S, pp, pr, sites, NULL, NULL);
            if (rc) break;
            if (compress) {
                cwrite_window(&cw, c->name, ws, C, sites);
            } else {
                for (uint32_t j = 0; j < C; j++) {
                    int n = orc_format_row(c->name, ws + j + 1, &sites[j], &sites[j + 1], row);
                    fwrite(row, 1, n, fo);
                }
            }
            tcall += now_sec() - t0;
            nsites += C;
            nwin++;
        }
    }
    if (!rc && compress) rc = cwrite_index(&cw, out);
    if (fo) fclose(fo);
    cwriter_close(&cw);
    if (stats) {
        stats->sites = nsites;
        stats->windows = nwin;
        stats->seconds_call = tcall;
    }
#undef ORC_CALLED
    if (wmask) {
        for (int ci = 0; ci < g->n; ci++) free(wmask[ci]);
        free(wmask);
    }
    free(sites), free(row), free(order), free(b0), free(b1), free(seen), free(idx);
    free(pm), free(S), free(pp), free(pr);
    orc_reads_free(r);
    orc_genome_free(g);
    return rc;
}

/* ======================================================== decompressor */

static int rd(FILE* f, void* p, size_t sz, size_t n) { return n == 0 || fread(p, sz, n, f) == n; }

static void dict_decode_i32(int32_t* out, uint64_t n, const uint32_t* words, uint64_t nw, const int32_t* dict,
                            uint64_t nd)
{
    if (n == 0 || nw == 0) return;
    int bits = orc_num_bits(nd);
    int per = (int)floor(32 / (double)bits);
    uint32_t mask = bits == 32 ? 0xffffffffu : ((1u << bits) - 1);
    for (uint64_t w = 0; w < nw; w++) {
        uint64_t m = (w == nw - 1) ? n - (uint64_t)per * w : (uint64_t)per;
        for (uint64_t e = 0; e < m; e++) out[w * per + e] = dict[(words[w] >> (bits * (m - e - 1))) & mask];
    }
}
static void dict_decode_u32(uint32_t* out, uint64_t n, const uint32_t* words, uint64_t nw, const uint32_t* dict,
                            uint64_t nd)
{
    dict_decode_i32((int32_t*)out, n, words, nw, (const int32_t*)dict, nd);
}
static void dict_decode_f64(double* out, uint64_t n, const uint32_t* words, uint64_t nw, const double* dict,
                            uint64_t nd)
{
    if (n == 0 || nw == 0) return;
    int bits = orc_num_bits(nd);
    int per = (int)floor(32 / (double)bits);
    uint32_t mask = bits == 32 ? 0xffffffffu : ((1u << bits) - 1);
    for (uint64_t w = 0; w < nw; w++) {
        uint64_t m = (w == nw - 1) ? n - (uint64_t)per * w : (uint64_t)per;
        for (uint64_t e = 0; e < m; e++) out[w * per + e] = dict[(words[w] >> (bits * (m - e - 1))) & mask];
    }
}

static int read_rle(FILE* f, int32_t* out_i, double* out_d, uint32_t C)
{
    uint64_t pa[6];
    if (!rd(f, pa, 8, 6)) return 0;
    uint64_t nr = pa[0];
    uint32_t* wv = (uint32_t*)malloc(4 * (pa[1] + 1));
    uint32_t* wl = (uint32_t*)malloc(4 * (pa[2] + 1));
    void* dv = malloc((out_d ? 8 : 4) * (pa[3] + 1));
    uint32_t* dl = (uint32_t*)malloc(4 * (pa[4] + 1));
    int ok = rd(f, wv, 4, pa[1]) && rd(f, wl, 4, pa[2]) && rd(f, dv, out_d ? 8 : 4, pa[3]) && rd(f, dl, 4, pa[4]);
    uint32_t* rl = (uint32_t*)malloc(4 * (nr + 1));
    dict_decode_u32(rl, nr, wl, pa[2], dl, pa[4]);
    uint64_t k = 0;
    if (out_d) {
        double* rv = (double*)malloc(8 * (nr + 1));
        dict_decode_f64(rv, nr, wv, pa[1], (double*)dv, pa[3]);
        for (uint64_t i = 0; i < nr; i++)
            for (uint32_t m = 0; m < rl[i] && k < C; m++) out_d[k++] = rv[i];
        free(rv);
    } else {
        int32_t* rv = (int32_t*)malloc(4 * (nr + 1));
        dict_decode_i32(rv, nr, wv, pa[1], (int32_t*)dv, pa[3]);
        for (uint64_t i = 0; i < nr; i++)
            for (uint32_t m = 0; m < rl[i] && k < C; m++) out_i[k++] = rv[i];
        free(rv);
    }
    free(wv), free(wl), free(dv), free(dl), free(rl);
    return ok && k == C;
}

/* cpp/decompression.cpp:130-255 decompressAll */
int orc_decompress(const char* prefix, const char* out)
{
    FILE* f12 = open_sfx(prefix, ".1_2", "r");
    if (!f12) return fail(ORC_E_IO, "cannot open %s.1_2", prefix);
    unsigned long nwin = 0;
    if (fscanf(f12, "%lu", &nwin) != 1) {
        fclose(f12);
        return fail(ORC_E_FORMAT, "bad .1_2");
    }
    FILE* f346 = open_sfx(prefix, ".3_4_6", "rb");
    FILE* f5 = open_sfx(prefix, ".5_7_8_9_14_16", "rb");
    FILE* f10 = open_sfx(prefix, ".10", "rb");
    FILE* f11 = open_sfx(prefix, ".11_12_13", "rb");
    FILE* f15 = open_sfx(prefix, ".15_17", "rb");
    FILE* fo = fopen(out, "w");
    int rc = (!f346 || !f5 || !f10 || !f11 || !f15 || !fo) ? fail(ORC_E_IO, "cannot open -C files") : ORC_OK;
    char name[1024];
    char* c3 = NULL;
    char *c4 = NULL, *c6 = NULL, *c10 = NULL;
    int32_t *c5 = NULL, *c7 = NULL, *c8 = NULL, *c9 = NULL, *c14 = NULL, *c11 = NULL, *c12 = NULL, *c13 = NULL,
            *c17 = NULL;
    double *c16 = NULL, *c15 = NULL;
    uint32_t capC = 0;
    for (unsigned long w = 0; !rc && w < nwin; w++) {
        unsigned long ps, C;
        if (fscanf(f12, "%1023s\t%lu\t%lu\n", name, &ps, &C) != 3) {
            rc = fail(ORC_E_FORMAT, "bad .1_2 row");
            break;
        }
        if (C + 1 > capC) {
            capC = (uint32_t)C + 1;
            c3 = (char*)realloc(c3, capC), c4 = (char*)realloc(c4, capC), c6 = (char*)realloc(c6, capC);
            c10 = (char*)realloc(c10, capC);
            c5 = (int32_t*)realloc(c5, 4 * capC), c7 = (int32_t*)realloc(c7, 4 * capC);
            c8 = (int32_t*)realloc(c8, 4 * capC), c9 = (int32_t*)realloc(c9, 4 * capC);
            c14 = (int32_t*)realloc(c14, 4 * capC), c11 = (int32_t*)realloc(c11, 4 * capC);
            c12 = (int32_t*)realloc(c12, 4 * capC), c13 = (int32_t*)realloc(c13, 4 * capC);
            c17 = (int32_t*)realloc(c17, 4 * capC);
            c16 = (double*)realloc(c16, 8 * capC), c15 = (double*)realloc(c15, 8 * capC);
        }
        /* .3_4_6 */
        uint64_t pa[7];
        if (!rd(f346, pa, 8, 7)) {
            rc = fail(ORC_E_FORMAT, "short .3_4_6");
            break;
        }
        if (pa[0] == 1) {
            for (unsigned long i = 0; i < C; i++) c3[i] = c4[i] = c6[i] = 'N';
        } else {
            uint64_t n4 = pa[5], n6 = pa[6], n3 = pa[4], nc = pa[3];
            int32_t *id4 = (int32_t*)malloc(4 * (n4 + 1)), *id6 = (int32_t*)malloc(4 * (n6 + 1)),
                    *idn = (int32_t*)malloc(4 * (n3 + 1));
            char *v4 = (char*)malloc(n4 + 1), *v6 = (char*)malloc(n6 + 1);
            uint32_t* code = (uint32_t*)malloc(4 * (nc + 1));
            int ok = rd(f346, id4, 4, n4) && rd(f346, v4, 1, n4) && rd(f346, id6, 4, n6) && rd(f346, v6, 1, n6) &&
                     rd(f346, idn, 4, n3) && rd(f346, code, 4, nc);
            if (ok) {
                orc_decode_2bit(c3, C, code, nc);
                for (uint64_t i = 0; i < n3; i++) c3[idn[i]] = 'N';
                memcpy(c4, c3, C);
                for (uint64_t i = 0; i < n4; i++) c4[id4[i]] = v4[i];
                memcpy(c6, c3, C);
                for (uint64_t i = 0; i < n6; i++) c6[id6[i]] = v6[i];
            }
            free(id4), free(id6), free(idn), free(v4), free(v6), free(code);
            if (!ok) {
                rc = fail(ORC_E_FORMAT, "short .3_4_6 blob");
                break;
            }
        }
        /* .5_7_8_9_14_16 */
        int32_t nind;
        if (!rd(f5, &nind, 4, 1)) {
            rc = fail(ORC_E_FORMAT, "short .5_7");
            break;
        }
        if (nind == 1) {
            for (unsigned long i = 0; i < C; i++) c5[i] = c7[i] = c8[i] = c9[i] = c14[i] = 0, c16[i] = 255.0;
        } else if (!read_rle(f5, c5, NULL, C) || !read_rle(f5, c7, NULL, C) || !read_rle(f5, c8, NULL, C) ||
                   !read_rle(f5, c9, NULL, C) || !read_rle(f5, c14, NULL, C) || !read_rle(f5, NULL, c16, C)) {
            rc = fail(ORC_E_FORMAT, "bad RLE column");
            break;
        }
        /* .10 */
        if (!rd(f10, &nind, 4, 1)) {
            rc = fail(ORC_E_FORMAT, "short .10");
            break;
        }
        if (nind == 1) {
            for (unsigned long i = 0; i < C; i++) c10[i] = 'N';
        } else {
            uint64_t p3[3];
            if (!rd(f10, p3, 8, 3)) {
                rc = fail(ORC_E_FORMAT, "short .10");
                break;
            }
            uint32_t* code = (uint32_t*)malloc(4 * (p3[1] + 1));
            uint32_t* ni = (uint32_t*)malloc(4 * (p3[0] + 1));
            int ok = rd(f10, code, 4, p3[1]) && rd(f10, ni, 4, p3[0]);
            if (ok) {
                orc_decode_2bit(c10, C, code, p3[1]);
                for (uint64_t i = 0; i < p3[0]; i++) c10[ni[i]] = 'N';
            }
            free(code), free(ni);
            if (!ok) {
                rc = fail(ORC_E_FORMAT, "short .10 blob");
                break;
            }
        }
        /* .11_12_13 */
        if (!rd(f11, &nind, 4, 1)) {
            rc = fail(ORC_E_FORMAT, "short .11");
            break;
        }
        memset(c11, 0, 4 * C), memset(c12, 0, 4 * C), memset(c13, 0, 4 * C);
        if (nind != 1) {
            uint64_t p4[4];
            if (!rd(f11, p4, 8, 4)) {
                rc = fail(ORC_E_FORMAT, "short .11");
                break;
            }
            uint64_t nc = p4[1], n12 = p4[2], n13 = p4[3];
            int32_t *x12i = (int32_t*)malloc(4 * (n12 + 1)), *x12v = (int32_t*)malloc(4 * (n12 + 1));
            int32_t *x13i = (int32_t*)malloc(4 * (n13 + 1)), *x13v = (int32_t*)malloc(4 * (n13 + 1));
            int32_t* cid = (int32_t*)malloc(4 * (nc + 1));
            int32_t* vals[3];
            int ok = rd(f11, x12i, 4, n12) && rd(f11, x12v, 4, n12) && rd(f11, x13i, 4, n13) &&
                     rd(f11, x13v, 4, n13) && rd(f11, cid, 4, nc);
            for (int k = 0; k < 3; k++) {
                vals[k] = (int32_t*)malloc(4 * (nc + 1));
                uint64_t dp[2] = {0, 0};
                ok = ok && rd(f11, dp, 8, 2);
                uint32_t* words = (uint32_t*)malloc(4 * (dp[0] + 1));
                int32_t* dict = (int32_t*)malloc(4 * (dp[1] + 1));
                ok = ok && rd(f11, words, 4, dp[0]) && rd(f11, dict, 4, dp[1]);
                if (ok) dict_decode_i32(vals[k], nc, words, dp[0], dict, dp[1]);
                free(words), free(dict);
            }
            if (ok) {
                for (uint64_t i = 0; i < nc; i++) c11[cid[i]] = vals[0][i], c12[cid[i]] = vals[1][i], c13[cid[i]] = vals[2][i];
                for (uint64_t i = 0; i < n12; i++) c12[x12i[i]] = x12v[i];
                for (uint64_t i = 0; i < n13; i++) c13[x13i[i]] = x13v[i];
            }
            free(x12i), free(x12v), free(x13i), free(x13v), free(cid);
            for (int k = 0; k < 3; k++) free(vals[k]);
            if (!ok) {
                rc = fail(ORC_E_FORMAT, "short .11 blob");
                break;
            }
        }
        /* .15_17 */
        if (!rd(f15, &nind, 4, 1)) {
            rc = fail(ORC_E_FORMAT, "short .15");
            break;
        }
        for (unsigned long i = 0; i < C; i++) c15[i] = 1.0, c17[i] = 0;
        if (nind != 1) {
            int32_t n;
            if (!rd(f15, &n, 4, 1)) {
                rc = fail(ORC_E_FORMAT, "short .15");
                break;
            }
            int32_t* id = (int32_t*)malloc(4 * ((size_t)n + 1));
            double* val = (double*)malloc(8 * ((size_t)n + 1));
            int ok = rd(f15, id, 4, n) && rd(f15, val, 8, n);
            if (ok)
                for (int32_t i = 0; i < n; i++) c15[id[i]] = val[i];
            ok = ok && rd(f15, &n, 4, 1);
            if (ok) {
                id = (int32_t*)realloc(id, 4 * ((size_t)n + 1));
                ok = rd(f15, id, 4, n);
                if (ok)
                    for (int32_t i = 0; i < n; i++) c17[id[i]] = 1;
            }
            free(id), free(val);
            if (!ok) {
                rc = fail(ORC_E_FORMAT, "short .15 blob");
                break;
            }
        }
        for (unsigned long i = 0; i < C; i++) {
            if (c3[i] == 'N' && c4[i] == 'N' && c6[i] == 'N') {
                fprintf(fo, "%s\t%lu\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n", name, ps + i);
            } else {
                fprintf(fo, "%s\t%lu\t%c\t%c\t%d\t%c\t%d\t%d\t%d\t%c\t%d\t%d\t%d\t%d\t%#g\t%#g\t%d\n", name, ps + i,
                        c3[i], c4[i], c5[i], c6[i], c7[i], c8[i], c9[i], c10[i], c11[i], c12[i], c13[i], c14[i],
                        c15[i], c16[i], c17[i]);
            }
        }
    }
    fclose(f12);
    if (f346) fclose(f346);
    if (f5) fclose(f5);
    if (f10) fclose(f10);
    if (f11) fclose(f11);
    if (f15) fclose(f15);
    if (fo) fclose(fo);
    free(c3), free(c4), free(c6), free(c10), free(c5), free(c7), free(c8), free(c9), free(c14);
    free(c11), free(c12), free(c13), free(c17), free(c16), free(c15);
    return rc;
}
