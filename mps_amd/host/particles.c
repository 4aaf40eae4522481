/*
 * particles.c -- host particle container, particles.txt loader and the
 * dam-break scene generator.
 *
 * Container semantics follow mps_fluid3D_parallel/src/particle.c:11-58
 * (capacity-bounded add that reports and drops on overflow).  The loader
 * reads the format of mps_fluid3D_parallel/src/main.c:20-85 (2-D:
 * mps_fluid2D/src/main.c, five columns).  Unlike the reference's main(), the
 * caller sizes the container from the file (mps_count_particles), so scenes
 * above 100,000 particles load.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "mps_hip.h"

ParticleSystem *particle_system_create(int capacity)
{
    ParticleSystem *ps = (ParticleSystem *)calloc(1, sizeof *ps);
    if (!ps) return NULL;
    ps->particles = (Particle *)calloc(capacity > 0 ? (size_t)capacity : 1, sizeof(Particle));
    if (!ps->particles) {
        free(ps);
        return NULL;
    }
    ps->capacity = capacity;
    return ps;
}

void particle_system_free(ParticleSystem *ps)
{
    if (!ps) return;
    mps_hip_release(ps);
    free(ps->particles);
    free(ps);
}

int particle_system_add(ParticleSystem *ps, double pos[], double vel[], int type)
{
    if (ps->num >= ps->capacity) {
        fprintf(stderr, "Error: particle capacity exceeded (%d)\n", ps->capacity);
        return -1;
    }
    Particle *p = &ps->particles[ps->num];
    for (int d = 0; d < MPS_DIM; d++) {
        p->pos[d] = pos[d];
        p->vel[d] = vel[d];
        p->acc[d] = 0.0;
    }
    p->pressure = 0.0;
    p->n = 0.0;
    p->type = type;
    p->on_surface = 0;
    return ps->num++;
}

/* first non-blank character of a line, or NULL for blank/comment lines */
static const char *content(const char *line)
{
    while (*line == ' ' || *line == '\t') line++;
    if (*line == '#' || *line == '\n' || *line == '\r' || *line == '\0') return NULL;
    return line;
}

static int read_header(FILE *fp)
{
    char line[512];
    while (fgets(line, sizeof line, fp)) {
        const char *s = content(line);
        int n;
        if (s && sscanf(s, "%d", &n) == 1) return n;
    }
    return -1;
}

int mps_count_particles(const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int n = read_header(fp);
    fclose(fp);
    return n;
}

/* One fgets-sized piece of the file after the header, parsed. */
typedef struct {
    double pos[MPS_DIM], vel[MPS_DIM];
    int type;
    int blank;   /* blank or comment line */
    int got;     /* sscanf's return */
} RowParse;

#define LINE_CAP 512          /* the reference's fgets buffer */
#define LOAD_BATCH 65536      /* pieces parsed per parallel pass */

static void parse_piece(const char *text, size_t len, RowParse *r)
{
    char line[LINE_CAP];
    memcpy(line, text, len);
    line[len] = '\0';
    const char *s = content(line);
    r->blank = s == NULL;
    if (r->blank) return;
#if MPS_DIM == 3
    r->got = sscanf(s, "%lf %lf %lf %lf %lf %lf %d", &r->pos[0], &r->pos[1], &r->pos[2],
                    &r->vel[0], &r->vel[1], &r->vel[2], &r->type);
#else
    r->got = sscanf(s, "%lf %lf %lf %lf %d", &r->pos[0], &r->pos[1], &r->vel[0], &r->vel[1],
                    &r->type);
#endif
}

/* The rest of fp, NUL-terminated, or NULL where it does not fit in memory. */
static char *read_rest(FILE *fp, size_t *size)
{
    size_t cap = (size_t)1 << 20, n = 0;
    struct stat st;
    if (fstat(fileno(fp), &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0)
        cap = (size_t)st.st_size;   /* the whole file; the rest is no larger */
    char *buf = (char *)malloc(cap + 1);
    while (buf) {
        if (n == cap) {
            char *b = (char *)realloc(buf, 2 * cap + 1);
            if (!b) {
                free(buf);
                return NULL;
            }
            buf = b;
            cap *= 2;
        }
        size_t got = fread(buf + n, 1, cap - n, fp);
        n += got;
        if (got == 0) break;
    }
    if (buf) buf[n] = '\0';
    *size = n;
    return buf;
}

/*
 * The rows after the header are read whole and cut into the pieces the
 * reference's fgets loop sees (up to and including a newline, at most
 * LINE_CAP - 1 bytes).  Each batch of pieces is parsed in parallel with
 * the same sscanf call, then added in file order by one thread, which also
 * prints the warnings; the particles, counts and messages are those of the
 * line-by-line loop.  sscanf gets a short NUL-terminated copy of its piece,
 * since glibc's sscanf measures the whole string it is given.
 */
int mps_load_particles(const char *path, ParticleSystem *ps)
{
    FILE *fp = fopen(path, "r");
    if (!fp) {
        fprintf(stderr, "Error: cannot open particle file '%s'\n", path);
        return -1;
    }
    int expected = read_header(fp);
    if (expected <= 0) {
        fprintf(stderr, "Error: invalid particle count in '%s'\n", path);
        fclose(fp);
        return -1;
    }
#if MPS_DIM == 3
    const int want = 7;
#else
    const int want = 5;
#endif
    size_t size = 0;
    char *text = read_rest(fp, &size);
    fclose(fp);
    RowParse *rows = (RowParse *)malloc(sizeof(RowParse) * LOAD_BATCH);
    size_t *start = (size_t *)malloc(sizeof(size_t) * LOAD_BATCH);
    int *len = (int *)malloc(sizeof(int) * LOAD_BATCH);
    if (!text || !rows || !start || !len) {
        fprintf(stderr, "Error: out of memory reading '%s'\n", path);
        free(text);
        free(rows);
        free(start);
        free(len);
        return -1;
    }
    int loaded = 0, n_fluid = 0, n_wall = 0;
    size_t at = 0;
    while (loaded < expected && at < size) {
        int m = 0;
        while (m < LOAD_BATCH && at < size) {
            size_t lim = size - at < LINE_CAP - 1 ? size - at : LINE_CAP - 1;
            const char *nl = (const char *)memchr(text + at, '\n', lim);
            size_t l = nl ? (size_t)(nl - (text + at)) + 1 : lim;
            start[m] = at;
            len[m] = (int)l;
            m++;
            at += l;
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) parse_piece(text + start[i], (size_t)len[i], &rows[i]);
        for (int i = 0; i < m && loaded < expected; i++) {
            const RowParse *r = &rows[i];
            if (r->blank) continue;
            if (r->got != want) {
                char line[LINE_CAP];
                memcpy(line, text + start[i], (size_t)len[i]);
                line[len[i]] = '\0';
                fprintf(stderr, "Warning: skipping malformed line: %s", line);
                continue;
            }
            particle_system_add(ps, (double *)r->pos, (double *)r->vel, r->type);
            loaded++;
            if (r->type == FLUID_PARTICLE) n_fluid++;
            else if (r->type == WALL_PARTICLE) n_wall++;
        }
    }
    free(text);
    free(rows);
    free(start);
    free(len);
    if (loaded != expected)
        fprintf(stderr, "Warning: expected %d particles but loaded %d\n", expected, loaded);
    printf("Loaded particles: %d fluid, %d wall, %d total\n", n_fluid, n_wall, ps->num);
    return 0;
}

/*
 * Lattice typing of one site (gen_dambreak.py:127-158): wall where any
 * index is on or beyond a tank face, dummy where it is wall_layers or more
 * beyond, fluid inside the column; -1 = empty site.
 */
static int site_type(const long idx[3], const long face_hi[3], int dim, int wl,
                     const double pos[3], const double water[3])
{
    int t = -1;
    for (int d = 0; d < dim; d++) {
        if (idx[d] <= 0) t = WALL_PARTICLE;
        if (d != 1 && idx[d] >= face_hi[d]) t = WALL_PARTICLE; /* no lid on y */
    }
    for (int d = 0; d < dim; d++) {
        if (idx[d] <= -wl) t = DUMMY_PARTICLE;
        if (d != 1 && idx[d] >= face_hi[d] + wl) t = DUMMY_PARTICLE;
    }
    if (t < 0) {
        const double eps = 1.0e-10;
        int inside = 1;
        for (int d = 0; d < dim; d++)
            if (!(idx[d] >= 1 && pos[d] <= water[d] + eps)) inside = 0;
        if (inside) t = FLUID_PARTICLE;
    }
    return t;
}

int mps_gen_dambreak(const double dims[6], double l0, int wall_layers,
                     int dummy_layers, ParticleSystem *ps, const char *path)
{
    const int dim = MPS_DIM;
    const int layers = wall_layers + dummy_layers;
    const double domain[3] = { dims[0], dims[1], dims[2] };
    const double water[3] = { dims[3], dims[4], dims[5] };
    long lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 }, face_hi[3] = { 0, 0, 0 };
    for (int d = 0; d < dim; d++) {
        /* 3-D truncates span/l0 (gen_dambreak.py:106-112); 2-D rounds
         * (mps_fluid2D/tools/gen_dambreak.py:109-114) */
        double q = domain[d] / l0;
        long cells = (long)q;
        if (dim == 2) {
            /* Python round() of a positive value: half to even */
            double frac = q - (double)cells;
            if (frac > 0.5 || (frac == 0.5 && (cells & 1))) cells += 1;
        }
        face_hi[d] = cells;
        lo[d] = -(layers - 1);
        hi[d] = (d == 1) ? cells : cells + (layers - 1);
    }

    FILE *fp = NULL;
    long count = 0, n_fluid = 0, n_wall = 0, n_dummy = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            if (path) {
                fp = fopen(path, "w");
                if (!fp) return -1;
                fprintf(fp, "# MPS %dD Dam Break - Initial Particle Configuration\n", dim);
                fprintf(fp, "# Fluid: %ld, Wall: %ld, Dummy: %ld, Total: %ld\n",
                        n_fluid, n_wall, n_dummy, count);
                fprintf(fp, "%ld\n", count);
            }
            if (!ps && !fp) break;
        }
        long idx[3] = { 0, 0, 0 };
        for (idx[0] = lo[0]; idx[0] <= hi[0]; idx[0]++)
        for (idx[1] = lo[1]; idx[1] <= hi[1]; idx[1]++)
        for (idx[2] = (dim == 3 ? lo[2] : 0); idx[2] <= (dim == 3 ? hi[2] : 0); idx[2]++) {
            double pos[3];
            for (int d = 0; d < 3; d++) pos[d] = (double)idx[d] * l0;
            int t = site_type(idx, face_hi, dim, wall_layers, pos, water);
            if (t < 0) continue;
            if (pass == 0) {
                count++;
                if (t == FLUID_PARTICLE) n_fluid++;
                else if (t == WALL_PARTICLE) n_wall++;
                else n_dummy++;
                continue;
            }
            /* round-trip through the %.8e text the file carries */
            char buf[64];
            double v[MPS_DIM], x[MPS_DIM];
            for (int d = 0; d < dim; d++) {
                snprintf(buf, sizeof buf, "%.8e", pos[d]);
                x[d] = strtod(buf, NULL);
                v[d] = 0.0;
            }
            if (fp) {
                for (int d = 0; d < dim; d++) fprintf(fp, "%.8e ", pos[d]);
                for (int d = 0; d < dim; d++) fprintf(fp, "%.8e ", 0.0);
                fprintf(fp, "%d\n", t);
            }
            if (ps && particle_system_add(ps, x, v, t) < 0) {
                if (fp) fclose(fp);
                return -1;
            }
        }
    }
    if (fp) fclose(fp);
    return (int)count;
}
