/*
 * elp_stats.c -- polymer statistics of ELP configurations (see elp_stats.h).
 * Restates the reference analysis formulas; each function cites its source.
 */
#include "elp_stats.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct elp_series {
    int ring;
    int L[3];
    int n_chain, n_mono;
    int max_frames, n_frames;
    int32_t* pos;        /* [frame][chain][mono][3] unwrapped t,u,v; NULL in summary mode */
    /* summary mode, per frame and chain (unwrapped): */
    int32_t* first;      /* [frame][chain][3] monomer 0 */
    int64_t* sum;        /* [frame][chain][3] sum over monomers */
    int32_t* mid;        /* [frame][chain][3] monomer polSize/2 */
    double* rg2s;        /* [frame][chain] */
    int32_t* nsls;       /* [frame][chain] stored-length bonds */
    double* cnts;        /* [frame][n_mono] contact pairs by separation */
};

static inline int32_t* frame_chain(const elp_series* s, int f, int c) {
    return s->pos + (((size_t)f * s->n_chain + c) * s->n_mono) * 3;
}

elp_series* elp_series_new(int ring, int LT, int LU, int LV, int n_chain, int n_mono, int max_frames) {
    if (LT <= 0 || LU <= 0 || LV <= 0 || n_chain <= 0 || n_mono < 2 || max_frames <= 0) return NULL;
    elp_series* s = (elp_series*)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->ring = ring ? 1 : 0;
    s->L[0] = LT; s->L[1] = LU; s->L[2] = LV;
    s->n_chain = n_chain; s->n_mono = n_mono; s->max_frames = max_frames;
    s->pos = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)max_frames * n_chain * n_mono);
    if (!s->pos) { free(s); return NULL; }
    return s;
}

elp_series* elp_series_new_summary(int ring, int LT, int LU, int LV, int n_chain, int n_mono, int max_frames) {
    if (LT <= 0 || LU <= 0 || LV <= 0 || n_chain <= 0 || n_mono < 2 || max_frames <= 0) return NULL;
    elp_series* s = (elp_series*)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->ring = ring ? 1 : 0;
    s->L[0] = LT; s->L[1] = LU; s->L[2] = LV;
    s->n_chain = n_chain; s->n_mono = n_mono; s->max_frames = max_frames;
    const size_t fc = (size_t)max_frames * n_chain;
    s->first = (int32_t*)malloc(sizeof(int32_t) * 3 * fc);
    s->sum = (int64_t*)malloc(sizeof(int64_t) * 3 * fc);
    s->mid = (int32_t*)malloc(sizeof(int32_t) * 3 * fc);
    s->rg2s = (double*)malloc(sizeof(double) * fc);
    s->nsls = (int32_t*)malloc(sizeof(int32_t) * fc);
    s->cnts = (double*)malloc(sizeof(double) * (size_t)max_frames * n_mono);
    if (!s->first || !s->sum || !s->mid || !s->rg2s || !s->nsls || !s->cnts) { elp_series_free(s); return NULL; }
    return s;
}

void elp_series_free(elp_series* s) {
    if (!s) return;
    free(s->pos);
    free(s->first); free(s->sum); free(s->mid); free(s->rg2s); free(s->nsls); free(s->cnts);
    free(s);
}

int elp_series_frames(const elp_series* s) { return s ? s->n_frames : 0; }

/* Bond code -> (dt,du,dv) (gpupol2/ringpol3.cu:58-70; analysis/mode_io.c:547-550). */
static inline void bond_delta(int b, int d[3]) {
    const int w = (b >> 3) & 1;
    d[0] = (b & 1) - w;
    d[1] = ((b >> 1) & 1) - w;
    d[2] = ((b >> 2) & 1) - w;
}

static inline int imin_image(int x, int L) {
    /* shift x by multiples of L into (-L/2, L/2] as the while loops of
       analysis/mode_io.c:532-537 do for a difference */
    while (x > L / 2) x -= L;
    while (x < -L / 2) x += L;
    return x;
}

/* Unwrap one chain of frame f (analysis/mode_io.c:531-581). */
static void place_chain(elp_series* s, int f, int c, const int32_t start[3], const uint8_t* bonds) {
    int32_t* p = frame_chain(s, f, c);
    const int n = s->n_mono;
    int st[3] = {start[0], start[1], start[2]};
    if (f > 0) {
        const int32_t* q = frame_chain(s, f - 1, c);
        for (int k = 0; k < 3; k++) st[k] = q[k] + imin_image(st[k] - q[k], s->L[k]);
    }
    for (int k = 0; k < 3; k++) p[k] = st[k];
    for (int i = 0; i < n - 1; i++) {
        int d[3];
        bond_delta(bonds[i] & 0xf, d);
        for (int k = 0; k < 3; k++) p[3 * (i + 1) + k] = p[3 * i + k] + d[k];
    }
    if (f > 0) {
        /* whole-box shift so the mean position moves by less than half a box */
        const int32_t* q = frame_chain(s, f - 1, c);
        for (int k = 0; k < 3; k++) {
            double a = 0, b = 0;
            for (int i = 0; i < n; i++) { a += p[3 * i + k]; b += q[3 * i + k]; }
            a /= n; b /= n;
            int sh = 0;
            while (a - b + sh > s->L[k] / 2) sh -= s->L[k];
            while (a - b + sh < -s->L[k] / 2) sh += s->L[k];
            if (sh) for (int i = 0; i < n; i++) p[3 * i + k] += sh;
        }
    }
}

int elp_series_add(elp_series* s, const int32_t* starts, const uint8_t* bonds, int stride) {
    if (!s || !s->pos || !starts || !bonds || stride < s->n_mono - 1) return -1;
    if (s->n_frames >= s->max_frames) return -1;
    for (int c = 0; c < s->n_chain; c++) place_chain(s, s->n_frames, c, starts + 3 * c, bonds + (size_t)c * stride);
    s->n_frames++;
    return 0;
}

/* Summary-mode frame: the unwrapping of place_chain done on the per-chain
   reductions (start against the previous frame's monomer 0 by the minimum
   image, then whole boxes so that the mean moves by less than half a box;
   the mean is sum/n computed from the exact integer sum, as place_chain's
   double accumulation of integer positions gives it). */
int elp_series_add_summary(elp_series* s, const int32_t* start3, const int64_t* sum3, const int32_t* mid3,
                           const double* rg2, const int32_t* nsl, const uint64_t* contacts) {
    if (!s || s->pos || !start3 || !sum3 || !mid3 || !rg2 || !nsl || !contacts) return -1;
    if (s->n_frames >= s->max_frames) return -1;
    const int f = s->n_frames, nc = s->n_chain, n = s->n_mono;
    for (int c = 0; c < nc; c++) {
        const size_t at = (size_t)f * nc + c, prev = at - (size_t)nc;
        for (int k = 0; k < 3; k++) {
            int st = start3[3 * c + k];
            if (f > 0) {
                const int q0 = s->first[3 * prev + k];
                st = q0 + imin_image(st - q0, s->L[k]);
            }
            int64_t S = (int64_t)n * st + sum3[3 * c + k];
            int sh = 0;
            if (f > 0) {
                const double a = (double)S / n, b = (double)s->sum[3 * prev + k] / n;
                while (a - b + sh > s->L[k] / 2) sh -= s->L[k];
                while (a - b + sh < -s->L[k] / 2) sh += s->L[k];
            }
            s->first[3 * at + k] = st + sh;
            s->sum[3 * at + k] = S + (int64_t)n * sh;
            s->mid[3 * at + k] = st + mid3[3 * c + k] + sh;
        }
        s->rg2s[at] = rg2[c];
        s->nsls[at] = nsl[c];
    }
    for (int g = 0; g < n; g++) s->cnts[(size_t)f * n + g] = (double)contacts[g];
    s->n_frames++;
    return 0;
}

static int hexval(int ch) {
    if (ch >= '0' && ch <= '9') return ch - '0';
    if (ch >= 'a' && ch <= 'f') return ch - 'a' + 10;
    if (ch >= 'A' && ch <= 'F') return ch - 'A' + 10;
    return -1;
}

/* tuv text: "LT= L\nLU= L\nLV= L\nnp= n\nmaxPolLength= M\n" then per chain
   "len= N\n t u v\t<hex x N>\n" (gpupol2/gpupol_io.c:105-203). */
int elp_series_add_tuv(elp_series* s, const char* path) {
    if (!s || !s->pos) return -1;
    if (s->n_frames >= s->max_frames) return -1;
    FILE* fp = fopen(path, "r");
    if (!fp) return -2;
    int L[3], np, maxlen;
    int rc = 0;
    if (fscanf(fp, "%*s %d %*s %d %*s %d %*s %d %*s %d", &L[0], &L[1], &L[2], &np, &maxlen) != 5) { fclose(fp); return -2; }
    if (L[0] != s->L[0] || L[1] != s->L[1] || L[2] != s->L[2] || np != s->n_chain) { fclose(fp); return -3; }
    char* txt = (char*)malloc((size_t)maxlen + 16);
    uint8_t* b = (uint8_t*)malloc((size_t)s->n_mono + 1);
    char fmt[32];
    snprintf(fmt, sizeof fmt, "%%%ds", maxlen + 8);
    for (int c = 0; c < np && !rc; c++) {
        int len, st[3];
        if (fscanf(fp, "%*s %d %d %d %d", &len, &st[0], &st[1], &st[2]) != 4 || fscanf(fp, fmt, txt) != 1) { rc = -2; break; }
        if (len != s->n_mono || (int)strlen(txt) != len) { rc = -3; break; }
        for (int i = 0; i < len; i++) {
            const int h = hexval(txt[i]);
            if (h < 0) { rc = -2; break; }
            b[i] = (uint8_t)h;
        }
        if (!rc) {
            const int32_t s3[3] = {st[0], st[1], st[2]};
            place_chain(s, s->n_frames, c, s3, b);
        }
    }
    free(txt);
    free(b);
    fclose(fp);
    if (!rc) s->n_frames++;
    return rc;
}

int elp_series_positions(const elp_series* s, int f, int c, int32_t* out) {
    if (!s || !s->pos || f < 0 || f >= s->n_frames || c < 0 || c >= s->n_chain || !out) return -1;
    memcpy(out, frame_chain(s, f, c), sizeof(int32_t) * 3 * (size_t)s->n_mono);
    return 0;
}

static inline void xyz_of(const int32_t* p, double* x, double* y, double* z) {
    *x = (double)(p[0] + p[1] - p[2]);
    *y = (double)(p[0] - p[1]);
    *z = (double)p[2];
}

/* centre of mass in bond units (analysis/mode_compute.c:205-217) */
static void cms_of(const elp_series* s, const int32_t* p, double c[3]) {
    c[0] = c[1] = c[2] = 0;
    for (int i = 0; i < s->n_mono; i++) {
        double x, y, z;
        xyz_of(p + 3 * i, &x, &y, &z);
        c[0] += x; c[1] += y; c[2] += z;
    }
    const double f = 1.0 / (s->n_mono * sqrt(2.0));
    c[0] *= f; c[1] *= f; c[2] *= f;
}

/* analysis/mode_compute.c:191-203 */
static double rg2_of(const elp_series* s, const int32_t* p) {
    double c[3];
    cms_of(s, p, c);
    const double r = 1.0 / sqrt(2.0);
    double acc = 0;
    for (int i = 0; i < s->n_mono; i++) {
        double x, y, z;
        xyz_of(p + 3 * i, &x, &y, &z);
        const double dx = x * r - c[0], dy = y * r - c[1], dz = z * r - c[2];
        acc += dx * dx + dy * dy + dz * dz;
    }
    return acc / s->n_mono;
}

int elp_series_rg2(const elp_series* s, int f, double* rg2) {
    if (!s || f < 0 || f >= s->n_frames || !rg2) return -1;
    for (int c = 0; c < s->n_chain; c++)
        rg2[c] = s->pos ? rg2_of(s, frame_chain(s, f, c)) : s->rg2s[(size_t)f * s->n_chain + c];
    return 0;
}

/* number of bonds of the chain: polSize (analysis/mode_io.c:596-603) */
static inline int pol_size(const elp_series* s) { return s->ring ? s->n_mono : s->n_mono - 1; }

/* 12 FCC neighbour offsets in (t,u,v): the displacements of the valid bond codes */
static int g_nb[12][3];
static int g_nb_init = 0;
static void init_nb(void) {
    if (g_nb_init) return;
    int k = 0;
    for (int b = 1; b < 15; b++) {
        if (b == 3 || b == 12) continue;
        bond_delta(b, g_nb[k++]);
    }
    g_nb_init = 1;
}

int elp_series_statics(const elp_series* s, int n_therm, double* rg2, double* slrat, double* pc) {
    if (!s || n_therm < 0 || n_therm >= s->n_frames) return -1;
    init_nb();
    const int n = s->n_mono, ps = pol_size(s);
    const int neq = s->n_frames - n_therm;
    double srg = 0, ssl = 0;
    double* cnt = pc ? (double*)calloc((size_t)n, sizeof(double)) : NULL;
    for (int f = n_therm; f < s->n_frames && !s->pos; f++) {
        for (int c = 0; c < s->n_chain; c++) {
            srg += s->rg2s[(size_t)f * s->n_chain + c];
            ssl += s->nsls[(size_t)f * s->n_chain + c] / (double)ps;
        }
        if (cnt)
            for (int g = 0; g < n; g++) cnt[g] += s->cnts[(size_t)f * n + g];
    }
    for (int f = n_therm; f < s->n_frames && s->pos; f++) {
        for (int c = 0; c < s->n_chain; c++) {
            const int32_t* p = frame_chain(s, f, c);
            srg += rg2_of(s, p);
            /* analysis/mode_compute.c:145-153 */
            int nsl = 0;
            for (int i = 0; i < ps; i++) {
                const int32_t* a = p + 3 * i;
                const int32_t* b = p + 3 * ((i + 1) % n);
                nsl += (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]);
            }
            ssl += nsl / (double)ps;
            /* analysis/mode_compute.c:429-504: pairs on one site or on
               nearest-neighbour sites, in the chain's own (unwrapped) frame */
            if (cnt) {
                for (int i = 0; i < n; i++) {
                    const int32_t* a = p + 3 * i;
                    for (int j = i + 1; j < n; j++) {
                        const int32_t* b = p + 3 * j;
                        const int d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
                        if (d0 < -1 || d0 > 1 || d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) continue;
                        int hit = (d0 == 0 && d1 == 0 && d2 == 0);
                        for (int k = 0; k < 12 && !hit; k++) hit = (d0 == g_nb[k][0] && d1 == g_nb[k][1] && d2 == g_nb[k][2]);
                        if (hit) cnt[j - i] += 1.0;
                    }
                }
            }
        }
    }
    const double norm = (double)s->n_chain * neq;
    if (rg2) *rg2 = srg / norm;
    if (slrat) *slrat = ssl / norm;
    if (cnt) {
        /* analysis/mode_io.c:54-62 */
        for (int g = 0; g < n; g++) pc[g] = cnt[g] / (norm * (n - g));
        free(cnt);
    }
    return 0;
}

int elp_series_summary(const elp_series* s, int f, int32_t* start3, int64_t* sum3, int32_t* mid3, double* rg2,
                       int32_t* nsl, uint64_t* contacts) {
    if (!s || !s->pos || f < 0 || f >= s->n_frames || !start3 || !sum3 || !mid3 || !rg2 || !nsl || !contacts) return -1;
    init_nb();
    const int n = s->n_mono, ps = pol_size(s), mono = ps / 2;
    for (int g = 0; g < n; g++) contacts[g] = 0;
    for (int c = 0; c < s->n_chain; c++) {
        const int32_t* p = frame_chain(s, f, c);
        for (int k = 0; k < 3; k++) {
            const int L = s->L[k];
            start3[3 * c + k] = ((p[k] % L) + L) % L;
            int64_t S = 0;
            for (int i = 0; i < n; i++) S += p[3 * i + k] - p[k];
            sum3[3 * c + k] = S;
            mid3[3 * c + k] = p[3 * mono + k] - p[k];
        }
        rg2[c] = rg2_of(s, p);
        int z = 0;
        for (int i = 0; i < ps; i++) {
            const int32_t* a = p + 3 * i;
            const int32_t* b = p + 3 * ((i + 1) % n);
            z += (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]);
        }
        nsl[c] = z;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                const int d0 = p[3 * j] - p[3 * i], d1 = p[3 * j + 1] - p[3 * i + 1], d2 = p[3 * j + 2] - p[3 * i + 2];
                if (d0 < -1 || d0 > 1 || d1 < -1 || d1 > 1 || d2 < -1 || d2 > 1) continue;
                int hit = (d0 == 0 && d1 == 0 && d2 == 0);
                for (int k = 0; k < 12 && !hit; k++) hit = (d0 == g_nb[k][0] && d1 == g_nb[k][1] && d2 == g_nb[k][2]);
                contacts[j - i] += (uint64_t)hit;
            }
    }
    return 0;
}

/* w1, w3: the share of the system mean's displacement taken off g1 and g3 */
static int series_msd(const elp_series* s, int n_therm, int n_out, double* g1, double* g3, double w1, double w3) {
    if (!s || n_therm < 0 || n_therm >= s->n_frames || n_out <= 0) return -1;
    const int nf = s->n_frames, nc = s->n_chain;
    const int mono = pol_size(s) / 2;
    const double r = 1.0 / sqrt(2.0);
    /* per frame: chain centres of mass and the system mean (bond units) */
    double* cms = (double*)malloc(sizeof(double) * 3 * (size_t)nf * nc);
    double* sys = (double*)calloc(3 * (size_t)nf, sizeof(double));
    for (int f = 0; f < nf; f++)
        for (int c = 0; c < nc; c++) {
            double* cc = cms + 3 * ((size_t)f * nc + c);
            if (s->pos) {
                cms_of(s, frame_chain(s, f, c), cc);
            } else {
                /* the sum of x = t+u-v over monomers, as cms_of accumulates it */
                const int64_t* q = s->sum + 3 * ((size_t)f * nc + c);
                const double fac = 1.0 / (s->n_mono * sqrt(2.0));
                cc[0] = (double)(q[0] + q[1] - q[2]) * fac;
                cc[1] = (double)(q[0] - q[1]) * fac;
                cc[2] = (double)q[2] * fac;
            }
            for (int k = 0; k < 3; k++) sys[3 * f + k] += cc[k] / nc;
        }
    for (int k = 0; k < n_out; k++) {
        double a1 = 0, a3 = 0;
        long num = 0;
        for (int t = n_therm; t + k < nf; t++) {
            const int t2 = t + k;
            const double dr[3] = {sys[3 * t] - sys[3 * t2], sys[3 * t + 1] - sys[3 * t2 + 1], sys[3 * t + 2] - sys[3 * t2 + 2]};
            for (int c = 0; c < nc; c++) {
                double x1, y1, z1, x2, y2, z2;
                xyz_of(s->pos ? frame_chain(s, t, c) + 3 * mono : s->mid + 3 * ((size_t)t * nc + c), &x1, &y1, &z1);
                xyz_of(s->pos ? frame_chain(s, t2, c) + 3 * mono : s->mid + 3 * ((size_t)t2 * nc + c), &x2, &y2, &z2);
                const double mx = (x1 - x2) * r - w1 * dr[0], my = (y1 - y2) * r - w1 * dr[1],
                             mz = (z1 - z2) * r - w1 * dr[2];
                a1 += mx * mx + my * my + mz * mz;
                const double* c1 = cms + 3 * ((size_t)t * nc + c);
                const double* c2 = cms + 3 * ((size_t)t2 * nc + c);
                const double cx = c1[0] - c2[0] - w3 * dr[0], cy = c1[1] - c2[1] - w3 * dr[1],
                             cz = c1[2] - c2[2] - w3 * dr[2];
                a3 += cx * cx + cy * cy + cz * cz;
                num++;
            }
        }
        if (g1) g1[k] = num ? a1 / num : NAN;
        if (g3) g3[k] = num ? a3 / num : NAN;
    }
    free(cms);
    free(sys);
    return 0;
}

int elp_series_msd(const elp_series* s, int n_therm, int n_out, double* g1, double* g3) {
    return series_msd(s, n_therm, n_out, g1, g3, 1.0, 1.0);
}

/* analysis/mode_compute.c:506-570 with the drift of analysis/mode_compute.c:244-251 and
   analysis/mode_io.c:640-652: avg_pos holds the mean over chains of cms / sqrt(2); g3 subtracts it
   from a centre-of-mass displacement in bond units, g1 from a middle-monomer displacement in
   lattice units before dividing by 2 */
int elp_series_msd_reference(const elp_series* s, int n_therm, int n_out, double* g1, double* g3) {
    return series_msd(s, n_therm, n_out, g1, g3, 0.5, 1.0 / sqrt(2.0));
}
