/*
 * io.c -- snapshot writers with the reference's exact text formats:
 * CSV "x,y,z,vx,vy,vz,pressure,type" at %.8e (mps_fluid3D_parallel/src/io.c:10-32;
 * 2-D "x,y,vx,vy,pressure,type", mps_fluid2D/src/io.c:10-32) and legacy
 * ASCII VTK (mps_fluid3D_parallel/src/io.c:37-107; 2-D writes z as "0.0"
 * and titles the file "MPS Simulation Step").  Ghost particles are skipped.
 *
 * At 10^6 particles a snapshot is ~100 MB of text, so rows are formatted in
 * parallel (OpenMP) into per-chunk buffers that are then written in order;
 * the bytes are identical to a serial fprintf loop.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mps_hip.h"

typedef int (*row_fmt)(char *dst, size_t cap, const Particle *p);

static int fmt_csv(char *dst, size_t cap, const Particle *p)
{
#if MPS_DIM == 3
    return snprintf(dst, cap, "%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%d\n",
                    p->pos[0], p->pos[1], p->pos[2], p->vel[0], p->vel[1], p->vel[2],
                    p->pressure, p->type);
#else
    return snprintf(dst, cap, "%.8e,%.8e,%.8e,%.8e,%.8e,%d\n",
                    p->pos[0], p->pos[1], p->vel[0], p->vel[1], p->pressure, p->type);
#endif
}

static int fmt_point(char *dst, size_t cap, const Particle *p)
{
#if MPS_DIM == 3
    return snprintf(dst, cap, "%.8e %.8e %.8e\n", p->pos[0], p->pos[1], p->pos[2]);
#else
    return snprintf(dst, cap, "%.8e %.8e 0.0\n", p->pos[0], p->pos[1]);
#endif
}

static int fmt_pressure(char *dst, size_t cap, const Particle *p)
{
    return snprintf(dst, cap, "%.8e\n", p->pressure);
}

static int fmt_type(char *dst, size_t cap, const Particle *p)
{
    return snprintf(dst, cap, "%d\n", p->type);
}

static int fmt_velocity(char *dst, size_t cap, const Particle *p)
{
#if MPS_DIM == 3
    return snprintf(dst, cap, "%.8e %.8e %.8e\n", p->vel[0], p->vel[1], p->vel[2]);
#else
    return snprintf(dst, cap, "%.8e %.8e 0.0\n", p->vel[0], p->vel[1]);
#endif
}

/* Writes fmt(p) for every non-ghost particle, in index order. */
static void write_rows(FILE *fp, const ParticleSystem *ps, row_fmt fmt)
{
    const int chunk = 8192;
    const int n = ps->num;
    const int nchunks = (n + chunk - 1) / chunk;
    const size_t row_cap = 192;
    char *buf = (char *)malloc((size_t)chunk * row_cap * (nchunks > 0 ? (size_t)(nchunks < 64 ? nchunks : 64) : 1));
    size_t *len = (size_t *)malloc(sizeof(size_t) * (nchunks < 64 ? (nchunks > 0 ? nchunks : 1) : 64));
    if (!buf || !len) {
        free(buf);
        free(len);
        char row[192];
        for (int i = 0; i < n; i++)
            if (ps->particles[i].type != GHOST_PARTICLE) {
                int k = fmt(row, sizeof row, &ps->particles[i]);
                fwrite(row, 1, (size_t)k, fp);
            }
        return;
    }
    /* process up to 64 chunks per batch so memory stays bounded */
    for (int c0 = 0; c0 < nchunks; c0 += 64) {
        int c1 = c0 + 64 < nchunks ? c0 + 64 : nchunks;
#pragma omp parallel for schedule(dynamic, 1)
        for (int c = c0; c < c1; c++) {
            char *dst = buf + (size_t)(c - c0) * chunk * row_cap;
            size_t used = 0;
            int i_end = (c + 1) * chunk < n ? (c + 1) * chunk : n;
            for (int i = c * chunk; i < i_end; i++) {
                if (ps->particles[i].type == GHOST_PARTICLE) continue;
                used += (size_t)fmt(dst + used, row_cap, &ps->particles[i]);
            }
            len[c - c0] = used;
        }
        for (int c = c0; c < c1; c++)
            fwrite(buf + (size_t)(c - c0) * chunk * row_cap, 1, len[c - c0], fp);
    }
    free(buf);
    free(len);
}

static FILE *open_snapshot(const char *dir, int step, const char *ext)
{
    char name[256];
    snprintf(name, sizeof name, "%s/output_%06d.%s", dir, step, ext);
    FILE *fp = fopen(name, "w");
    if (!fp) fprintf(stderr, "Error: cannot open %s\n", name);
    return fp;
}

void output_csv(const ParticleSystem *ps, int step, const char *output_dir)
{
    FILE *fp = open_snapshot(output_dir, step, "csv");
    if (!fp) return;
#if MPS_DIM == 3
    fputs("x,y,z,vx,vy,vz,pressure,type\n", fp);
#else
    fputs("x,y,vx,vy,pressure,type\n", fp);
#endif
    write_rows(fp, ps, fmt_csv);
    fclose(fp);
}

void output_vtk(const ParticleSystem *ps, int step, const char *output_dir)
{
    int count = 0;
    for (int i = 0; i < ps->num; i++)
        if (ps->particles[i].type != GHOST_PARTICLE) count++;
    FILE *fp = open_snapshot(output_dir, step, "vtk");
    if (!fp) return;
    fputs("# vtk DataFile Version 3.0\n", fp);
#if MPS_DIM == 3
    fprintf(fp, "MPS 3D Simulation Step %d\n", step);
#else
    fprintf(fp, "MPS Simulation Step %d\n", step);
#endif
    fputs("ASCII\nDATASET UNSTRUCTURED_GRID\n", fp);
    fprintf(fp, "POINTS %d double\n", count);
    write_rows(fp, ps, fmt_point);
    fprintf(fp, "CELLS %d %d\n", count, count * 2);
    for (int k = 0; k < count; k++) fprintf(fp, "1 %d\n", k);
    fprintf(fp, "CELL_TYPES %d\n", count);
    for (int k = 0; k < count; k++) fputs("1\n", fp);
    fprintf(fp, "POINT_DATA %d\n", count);
    fputs("SCALARS pressure double 1\nLOOKUP_TABLE default\n", fp);
    write_rows(fp, ps, fmt_pressure);
    fputs("SCALARS type int 1\nLOOKUP_TABLE default\n", fp);
    write_rows(fp, ps, fmt_type);
    fputs("VECTORS velocity double\n", fp);
    write_rows(fp, ps, fmt_velocity);
    fclose(fp);
}
