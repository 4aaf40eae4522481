/*
 * mps_oracle.c -- CPU restatement of the reference MPS timestep.
 * TEST INFRASTRUCTURE ONLY (see mps_oracle.h): the product library never
 * links, loads or calls this file.
 *
 * Every function cites the reference lines it restates (paths relative to
 * the reference root; 3-D lines are mps_fluid3D_parallel/, 2-D lines
 * mps_fluid2D/).  Arithmetic is written in the reference's evaluation order
 * (C's left-to-right association) and must be compiled without FMA
 * contraction (-ffp-contract=off) to stay bit-identical.
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mps_oracle.h"

#define D MPS_DIM

/* ------------------------------------------------------------------ */
/* scalar helpers                                                      */
/* ------------------------------------------------------------------ */

/* kernel.c:12-20: w = re/r - 1, r clamped below at 0.01 l0, 0 beyond re */
double orc_weight(double r, double re, double l0)
{
    if (r >= re) return 0.0;
    const double r_min = 0.01 * l0;
    if (r < r_min) r = r_min;
    return re / r - 1.0;
}

/* The reference converts cell coordinates with C's (int) cast.  On x86-64
 * (cvttsd2si) NaN and values outside the int range become INT_MIN, which the
 * range check then rejects; in range it truncates toward zero. */
static int to_int_x86(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

/* r^2 = sum_d (x_j - x_i)^2 accumulated from 0 in axis order
 * (neighbor_search.c:214-219, operators.c:28-32) */
static double dist2(const Particle *pi, const Particle *pj, double dr[D])
{
    double r2 = 0.0;
    for (int d = 0; d < D; d++) {
        dr[d] = pj->pos[d] - pi->pos[d];
        r2 += dr[d] * dr[d];
    }
    return r2;
}

static double search_radius(const SimConfig *c)
{
    /* simulation.c:22-24 */
    double re = fmax(c->influence_radius_lap, c->influence_radius_n);
    if (c->surface_tension_enabled) re = fmax(re, c->influence_radius_st);
    return re;
}

/* ------------------------------------------------------------------ */
/* grid and binning                                                    */
/* ------------------------------------------------------------------ */

/* cell_list_create (neighbor_search.c:92-122) */
void orc_grid_make(orc_grid *g, double re, const double *dmin, const double *dmax)
{
    const double margin = 4.0 * re;
    g->size = re;
    g->total = 1;
    for (int d = 0; d < 3; d++) {
        if (d < D) {
            g->origin[d] = dmin[d] - margin;
            double span = (dmax[d] + margin) - g->origin[d];
            g->nc[d] = (int)(span / re) + 2;
        } else {
            g->origin[d] = 0.0;
            g->nc[d] = 1;
        }
        g->total *= g->nc[d];
    }
    g->inv = 1.0 / g->size;
}

/* neighbor_search.c:171-178: per-axis truncation, -1 when outside */
static int cell_coords(const orc_grid *g, const double *pos, int c[3])
{
    c[0] = c[1] = c[2] = 0;
    for (int d = 0; d < D; d++) {
        c[d] = to_int_x86((pos[d] - g->origin[d]) * g->inv);
        if (c[d] < 0 || c[d] >= g->nc[d]) return -1;
    }
    return 0;
}

long orc_cell_of(const orc_grid *g, const double *pos)
{
    int c[3];
    if (cell_coords(g, pos, c) < 0) return -1;
    /* 3-D (neighbor_search.c:180) (iz*ny + iy)*nx + ix; 2-D iy*nx + ix */
    return ((long)c[2] * g->nc[1] + c[1]) * g->nc[0] + c[0];
}

/* Members of every cell, ascending particle index inside a cell (a stable
 * counting sort).  Walking a cell backwards visits descending index, the
 * order the reference's prepend-linked lists give (neighbor_search.c:180-182). */
typedef struct {
    long *first;    /* [total+1] */
    int  *member;   /* [num] */
} bins;

static int bins_build(bins *b, const orc_grid *g, const ParticleSystem *ps, int include_ghosts)
{
    const int n = ps->num;
    b->first = (long *)calloc((size_t)g->total + 1, sizeof(long));
    b->member = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    long *cell = (long *)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
    if (!b->first || !b->member || !cell) {
        free(b->first); free(b->member); free(cell);
        return -1;
    }
    for (int i = 0; i < n; i++) {
        cell[i] = -1;
        if (!include_ghosts && ps->particles[i].type == GHOST_PARTICLE) continue;
        cell[i] = orc_cell_of(g, ps->particles[i].pos);
        if (cell[i] >= 0) b->first[cell[i] + 1]++;
    }
    for (long c = 0; c < g->total; c++) b->first[c + 1] += b->first[c];
    long *fill = (long *)malloc(sizeof(long) * (size_t)g->total);
    if (!fill) { free(b->first); free(b->member); free(cell); return -1; }
    memcpy(fill, b->first, sizeof(long) * (size_t)g->total);
    for (int i = 0; i < n; i++)
        if (cell[i] >= 0) b->member[fill[cell[i]]++] = i;
    free(fill);
    free(cell);
    return 0;
}

static void bins_free(bins *b)
{
    free(b->first);
    free(b->member);
}

/* clamped neighbour-cell range of neighbor_search.c:199-210 */
static void cell_range(const orc_grid *g, const int c[3], int lo[3], int hi[3])
{
    for (int d = 0; d < 3; d++) {
        lo[d] = c[d] > 0 ? c[d] - 1 : 0;
        hi[d] = c[d] < g->nc[d] - 1 ? c[d] + 1 : g->nc[d] - 1;
    }
}

/* ------------------------------------------------------------------ */
/* neighbour lists                                                     */
/* ------------------------------------------------------------------ */

int orc_nlist_reserve(orc_nlist *nl, int n, long cap)
{
    if (nl->start == NULL || nl->n < n) {
        free(nl->start);
        nl->start = (long *)calloc((size_t)n + 1, sizeof(long));
        if (!nl->start) return -1;
    }
    nl->n = n;
    if (cap > nl->cap) {
        free(nl->idx);
        nl->idx = (int *)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
        if (!nl->idx) { nl->cap = 0; return -1; }
        nl->cap = cap;
    }
    return 0;
}

void orc_nlist_free(orc_nlist *nl)
{
    free(nl->start);
    free(nl->idx);
    memset(nl, 0, sizeof *nl);
}

/* Visits the neighbours of i in reference order; with out==NULL only counts.
 * Returns the count (not truncated at max_nb). */
static int visit_neighbours(const ParticleSystem *ps, const orc_grid *g, const bins *b,
                            int i, double re2, int *out)
{
    const Particle *pi = &ps->particles[i];
    if (pi->type == GHOST_PARTICLE) return 0;
    int c[3];
    if (cell_coords(g, pi->pos, c) < 0) return 0;
    int lo[3], hi[3];
    cell_range(g, c, lo, hi);
    int cnt = 0;
    for (int cz = lo[2]; cz <= hi[2]; cz++)
        for (int cy = lo[1]; cy <= hi[1]; cy++)
            for (int cx = lo[0]; cx <= hi[0]; cx++) {
                long cell = ((long)cz * g->nc[1] + cy) * g->nc[0] + cx;
                for (long k = b->first[cell + 1] - 1; k >= b->first[cell]; k--) {
                    int j = b->member[k];
                    if (j == i) continue;
                    double dr[D];
                    if (dist2(pi, &ps->particles[j], dr) < re2) {
                        if (out) out[cnt] = j;
                        cnt++;
                    }
                }
            }
    return cnt;
}

/* neighbor_search_cell_linked_list (neighbor_search.c:146-240) */
int orc_search(const ParticleSystem *ps, const SimConfig *c, double re, orc_nlist *out)
{
    const int n = ps->num;
    orc_grid g;
    orc_grid_make(&g, re, c->domain_min, c->domain_max);
    bins b;
    if (bins_build(&b, &g, ps, 0) != 0) return -1;
    const double re2 = re * re;
    int *cnt = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!cnt) { bins_free(&b); return -1; }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) cnt[i] = visit_neighbours(ps, &g, &b, i, re2, NULL);
    int overflow = 0;
    long total = 0;
    for (int i = 0; i < n; i++) {
        if (!overflow && cnt[i] > c->max_neighbors) overflow = i + 1;
        total += cnt[i];
    }
    if (orc_nlist_reserve(out, n, total) != 0) { free(cnt); bins_free(&b); return -1; }
    out->start[0] = 0;
    for (int i = 0; i < n; i++) out->start[i + 1] = out->start[i] + cnt[i];
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) visit_neighbours(ps, &g, &b, i, re2, out->idx + out->start[i]);
    free(cnt);
    bins_free(&b);
    return overflow;
}

/* neighbor_search_brute_force (neighbor_search.c:44-81): all j != i, both
 * non-ghost, r^2 < re^2, ascending j */
int orc_search_brute(const ParticleSystem *ps, double re, orc_nlist *out)
{
    const int n = ps->num;
    const double re2 = re * re;
    long total = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            if (orc_nlist_reserve(out, n, total) != 0) return -1;
            out->start[0] = 0;
        }
        long pos = 0;
        for (int i = 0; i < n; i++) {
            const Particle *pi = &ps->particles[i];
            if (pi->type != GHOST_PARTICLE) {
                for (int j = 0; j < n; j++) {
                    if (j == i || ps->particles[j].type == GHOST_PARTICLE) continue;
                    double dr[D];
                    if (dist2(pi, &ps->particles[j], dr) < re2) {
                        if (pass == 1) out->idx[pos] = j;
                        pos++;
                    }
                }
            }
            if (pass == 1) out->start[i + 1] = pos;
        }
        total = pos;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* particle operators                                                  */
/* ------------------------------------------------------------------ */

/* calc_particle_number_density (operators.c:17-38) */
void orc_density(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl)
{
    const double re = c->influence_radius_n, l0 = c->particle_distance;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ps->num; i++) {
        double sum = 0.0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            double dr[D];
            double r2 = dist2(&ps->particles[i], &ps->particles[nl->idx[k]], dr);
            sum += orc_weight(sqrt(r2), re, l0);
        }
        ps->particles[i].n = sum;
    }
}

/* judge_free_surface (operators.c:206-219) */
void orc_surface(ParticleSystem *ps, double threshold)
{
    const double limit_scale = threshold;
    for (int i = 0; i < ps->num; i++) {
        Particle *p = &ps->particles[i];
        p->on_surface = (p->type == FLUID_PARTICLE) ? (p->n < limit_scale * ps->n0) : 0;
    }
}

/* calc_neighbor_count (operators.c:226-246) */
void orc_neighbor_count(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl)
{
    const double re2 = c->influence_radius_n * c->influence_radius_n;
    for (int i = 0; i < ps->num; i++) {
        int cnt = 0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            double dr[D];
            if (dist2(&ps->particles[i], &ps->particles[nl->idx[k]], dr) < re2) cnt++;
        }
        ps->particles[i].neighbor_count = cnt;
    }
}

/* judge_free_surface_by_count (operators.c:253-267) */
void orc_surface_by_count(ParticleSystem *ps, double beta)
{
    for (int i = 0; i < ps->num; i++) {
        Particle *p = &ps->particles[i];
        p->on_surface = (p->type == FLUID_PARTICLE) ? (p->neighbor_count < beta * ps->n0_count) : 0;
    }
}

/* Potential-type surface tension, liquid-liquid (operators.c:281-311) or
 * solid-liquid (:322-354): adds to acc of fluid i over partners of type
 * `partner` with r in [1e-9 l0, re_st). */
static void surface_tension(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl,
                            double cst, int partner)
{
    const double re_st = c->influence_radius_st, l0 = c->particle_distance;
#if MPS_DIM == 3
    const double coeff = cst / (c->density * l0 * l0 * l0);
#else
    /* the 2-D program scales by rho*l0^2 (mps_fluid2D/src/operators.c:
     * calc_surface_tension, calc_surface_tension_SL) */
    const double coeff = cst / (c->density * l0 * l0);
#endif
    for (int i = 0; i < ps->num; i++) {
        Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const Particle *pj = &ps->particles[nl->idx[k]];
            if (pj->type != partner) continue;
            double dr[D];
            double r = sqrt(dist2(pi, pj, dr));
            if (r < 1.0e-9 * l0 || r >= re_st) continue;
            double f = -(r - l0) * (r - re_st);
            for (int d = 0; d < D; d++) pi->acc[d] += coeff * f * (dr[d] / r);
        }
    }
}

#ifdef MPS_FRICTION_FIELDS
/* Hybrid wall friction of the 3-D serial program (mps_fluid3D/src/
 * operators.c:337-463), applied to u* before r* = r + dt u*
 * (mps_fluid3D/src/simulation.c:50-53), over the stale list.  A fluid
 * particle is on the contact line when its fluid-only number density is
 * below surface_threshold n0 and a wall particle lies within re_st; the
 * nearest such wall particle (first strict minimum in list order) gives
 * the wall normal from its own wall neighbours, the particle's fluid
 * neighbours give the free-surface normal, and their angle the contact
 * angle.  Slow particles (|u*| <= v_c) are held when that angle is within
 * delta_theta of the wetting angle; faster ones lose slip velocity along
 * the wall, at most all of it.  Each particle changes only its own
 * velocity and reads positions and types, so the order over i is free. */
static void friction(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl)
{
    const double re_st = c->influence_radius_st, re_n = c->influence_radius_n, l0 = c->particle_distance;
    const double theta_s = c->wetting_angle_SL * M_PI / 180.0;
    const double dth = c->friction_delta_theta * M_PI / 180.0;
    const double v_c = c->friction_velocity_threshold;
    const double dyn = c->dynamic_friction_alpha * c->viscosity * 2.0 * D / (ps->n0 * ps->lambda);
    for (int i = 0; i < ps->num; i++) {
        Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        double nf = 0.0, r2s = 1.0e30, dr[D];
        int js = -1;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const int j = nl->idx[k];
            const Particle *pj = &ps->particles[j];
            const double r2 = dist2(pi, pj, dr);
            if (pj->type == FLUID_PARTICLE) {
                if (r2 < re_n * re_n) nf += orc_weight(sqrt(r2), re_n, l0);
            } else if (pj->type == WALL_PARTICLE && r2 < re_st * re_st && r2 < r2s) {
                r2s = r2;
                js = j;
            }
        }
        if (js < 0 || nf >= c->surface_threshold * ps->n0) continue;
        /* wall normal: sum over the wall neighbours m of j* of
         * w(|r_j* - r_m|) (r_j* - r_m) / |r_j* - r_m|, normalised */
        const Particle *pw = &ps->particles[js];
        double nw[D], nlv[D], es[D];
        for (int d = 0; d < D; d++) nw[d] = nlv[d] = 0.0;
        for (long k = nl->start[js]; k < nl->start[js + 1]; k++) {
            const Particle *pm = &ps->particles[nl->idx[k]];
            if (pm->type != WALL_PARTICLE) continue;
            const double r = sqrt(dist2(pm, pw, dr));
            if (r < 1.0e-9 * l0) continue;
            const double w = orc_weight(r, re_st, l0);
            for (int d = 0; d < D; d++) nw[d] += w * dr[d] / r;
        }
        double m2 = 0.0;
        for (int d = 0; d < D; d++) m2 += nw[d] * nw[d];
        const double mw = sqrt(m2);
        if (mw < 1.0e-9) continue;
        for (int d = 0; d < D; d++) nw[d] /= mw;
        /* outward free-surface normal: minus the same sum over the fluid
         * neighbours of i */
        int nfl = 0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const Particle *pj = &ps->particles[nl->idx[k]];
            if (pj->type != FLUID_PARTICLE) continue;
            const double r = sqrt(dist2(pi, pj, dr));
            if (r < 1.0e-9 * l0) continue;
            const double w = orc_weight(r, re_st, l0);
            for (int d = 0; d < D; d++) nlv[d] -= w * dr[d] / r;
            nfl++;
        }
        if (nfl == 0) continue;
        m2 = 0.0;
        for (int d = 0; d < D; d++) m2 += nlv[d] * nlv[d];
        const double ml = sqrt(m2);
        if (ml < 1.0e-9) continue;
        for (int d = 0; d < D; d++) nlv[d] /= ml;
        /* contact angle, cos = -n_w . n_lv */
        double cs = 0.0;
        for (int d = 0; d < D; d++) cs -= nw[d] * nlv[d];
        const double theta = acos(fmax(-1.0, fmin(1.0, cs)));
        /* slip direction: n_lv projected on the wall plane, normalised */
        double dot = 0.0;
        for (int d = 0; d < D; d++) dot += nlv[d] * nw[d];
        m2 = 0.0;
        for (int d = 0; d < D; d++) {
            es[d] = nlv[d] - dot * nw[d];
            m2 += es[d] * es[d];
        }
        const double me = sqrt(m2);
        if (me < 1.0e-9) continue;
        for (int d = 0; d < D; d++) es[d] /= me;
        double v2 = 0.0;
        for (int d = 0; d < D; d++) v2 += pi->vel[d] * pi->vel[d];
        if (sqrt(v2) <= v_c) {
            if (fabs(theta - theta_s) <= dth)
                for (int d = 0; d < D; d++) pi->vel[d] = 0.0;
        } else {
            double sw = 0.0;
            for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
                const Particle *pj = &ps->particles[nl->idx[k]];
                if (pj->type != WALL_PARTICLE) continue;
                const double r = sqrt(dist2(pi, pj, dr));
                if (r < 1.0e-9 * l0) continue;
                sw += orc_weight(r, re_st, l0);
            }
            double vn = 0.0;
            for (int d = 0; d < D; d++) vn += pi->vel[d] * es[d];
            double dvn = -dyn * c->dt * sw * vn;
            if (dvn > vn) dvn = vn;
            if (dvn < -vn) dvn = -vn;
            for (int d = 0; d < D; d++) pi->vel[d] += dvn * es[d];
        }
    }
}

/* apply_friction as a call of its own (mps_fluid3D/include/operators.h:42):
 * the friction on ps's current velocities over the list of ps's current
 * positions.  Returns 0, or -1 when the list cannot be built. */
int orc_apply_friction(ParticleSystem *ps, const SimConfig *c)
{
    orc_nlist nl = {0};
    double re = fmax(c->influence_radius_lap, c->influence_radius_n);
    if (c->surface_tension_enabled) re = fmax(re, c->influence_radius_st);
    if (orc_search(ps, c, re, &nl)) return -1;
    if (c->friction_enabled) friction(ps, c, &nl);
    orc_nlist_free(&nl);
    return 0;
}
#endif

/* acc = gravity (fluid) or 0; viscosity over the stale list
 * (operators.c:50-85); surface tension if enabled; u* = u + dt acc;
 * hybrid friction in the 3-D serial layout; r* = r + dt u*
 * (simulation.c:31-67) */
void orc_visc_predict(ParticleSystem *ps, const SimConfig *c, const orc_nlist *stale)
{
    const int n = ps->num;
    const double re = c->influence_radius_lap, l0 = c->particle_distance;
    const double coeff = 2.0 * D / (ps->n0 * ps->lambda);
    const double nu = c->viscosity, dt = c->dt;
    for (int i = 0; i < n; i++)
        for (int d = 0; d < D; d++)
            ps->particles[i].acc[d] = (ps->particles[i].type == FLUID_PARTICLE) ? c->gravity[d] : 0.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        double lap[D];
        for (int d = 0; d < D; d++) lap[d] = 0.0;
        for (long k = stale->start[i]; k < stale->start[i + 1]; k++) {
            const Particle *pj = &ps->particles[stale->idx[k]];
            double dr[D];
            double w = orc_weight(sqrt(dist2(pi, pj, dr)), re, l0);
            for (int d = 0; d < D; d++) lap[d] += (pj->vel[d] - pi->vel[d]) * w;
        }
        for (int d = 0; d < D; d++) pi->acc[d] += nu * coeff * lap[d];
    }
    if (c->surface_tension_enabled) {
        surface_tension(ps, c, stale, ps->C_LL, FLUID_PARTICLE);
        if (ps->C_SL != 0.0) surface_tension(ps, c, stale, ps->C_SL, WALL_PARTICLE);
    }
    for (int i = 0; i < n; i++) {
        Particle *p = &ps->particles[i];
        if (p->type != FLUID_PARTICLE) continue;
        for (int d = 0; d < D; d++) p->vel[d] += dt * p->acc[d];
    }
#ifdef MPS_FRICTION_FIELDS
    if (c->friction_enabled) friction(ps, c, stale);
#endif
    for (int i = 0; i < n; i++) {
        Particle *p = &ps->particles[i];
        if (p->type != FLUID_PARTICLE) continue;
        for (int d = 0; d < D; d++) p->pos[d] += dt * p->vel[d];
    }
}

/* One partner of the collision loop (operators.c:405-427): updates vi. */
static void collide_with(const Particle *pi, const Particle *pj, double vi[D],
                         double cd2, double l0, double e)
{
    if (pj->type == GHOST_PARTICLE) return;
    double dr[D];
    double r2 = dist2(pi, pj, dr);
    if (r2 >= cd2) return;
    double r = sqrt(r2);
    if (r < 1.0e-12 * l0) return;
    double vn = 0.0;
    for (int d = 0; d < D; d++) vn += (vi[d] - pj->vel[d]) * (dr[d] / r);
    if (vn <= 0.0) return;
    double impulse = (1.0 + e) * 0.5 * vn;
    for (int d = 0; d < D; d++) vi[d] -= impulse * (dr[d] / r);
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

void orc_collision(ParticleSystem *ps, const SimConfig *c, const orc_nlist *stale)
{
    const int n = ps->num;
    const double l0 = c->particle_distance;
    const double cd = c->collision_distance_ratio * l0;
    const double cd2 = cd * cd;
    const double e = c->restitution_coeff, dt = c->dt;
    double *after = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * D);
    if (!after) return;
    for (int i = 0; i < n; i++)
        for (int d = 0; d < D; d++) after[(size_t)i * D + d] = ps->particles[i].vel[d];
#if MPS_DIM == 3
    /* operators.c:395-432: partners from the stale list */
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        const Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        double vi[D];
        for (int d = 0; d < D; d++) vi[d] = pi->vel[d];
        for (long k = stale->start[i]; k < stale->start[i + 1]; k++)
            collide_with(pi, &ps->particles[stale->idx[k]], vi, cd2, l0, e);
        for (int d = 0; d < D; d++) after[(size_t)i * D + d] = vi[d];
    }
#else
    /* mps_fluid2D/src/operators.c:341-379: every j in ascending order.  Only
     * partners with r^2 < cd^2 act, so the candidates are taken from a grid of
     * cell size >= cd (the 3x3 block around i) plus every particle outside the
     * grid, then visited in ascending index -- the same sequence of updates. */
    (void)stale;
    orc_grid g;
    orc_grid_make(&g, search_radius(c), c->domain_min, c->domain_max);
    bins b;
    if (bins_build(&b, &g, ps, 0) != 0) { free(after); return; }
    int n_out = 0;
    int *outside = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++)
        if (ps->particles[j].type != GHOST_PARTICLE && orc_cell_of(&g, ps->particles[j].pos) < 0)
            outside[n_out++] = j;
#pragma omp parallel
    {
        int cap = 256 + n_out;
        int *cand = (int *)malloc(sizeof(int) * (size_t)cap);
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            const Particle *pi = &ps->particles[i];
            if (pi->type != FLUID_PARTICLE) continue;
            double vi[D];
            for (int d = 0; d < D; d++) vi[d] = pi->vel[d];
            int cc[3];
            if (cell_coords(&g, pi->pos, cc) < 0) {
                for (int j = 0; j < n; j++)
                    if (j != i) collide_with(pi, &ps->particles[j], vi, cd2, l0, e);
            } else {
                int lo[3], hi[3], m = 0;
                cell_range(&g, cc, lo, hi);
                for (int cy = lo[1]; cy <= hi[1]; cy++)
                    for (int cx = lo[0]; cx <= hi[0]; cx++) {
                        long cell = (long)cy * g.nc[0] + cx;
                        for (long k = b.first[cell]; k < b.first[cell + 1]; k++) {
                            int j = b.member[k];
                            double dr[D];
                            if (j == i || dist2(pi, &ps->particles[j], dr) >= cd2) continue;
                            if (m == cap) { cap *= 2; cand = (int *)realloc(cand, sizeof(int) * (size_t)cap); }
                            cand[m++] = j;
                        }
                    }
                for (int q = 0; q < n_out; q++) {
                    if (outside[q] == i) continue;
                    if (m == cap) { cap *= 2; cand = (int *)realloc(cand, sizeof(int) * (size_t)cap); }
                    cand[m++] = outside[q];
                }
                qsort(cand, (size_t)m, sizeof(int), cmp_int);
                for (int q = 0; q < m; q++) collide_with(pi, &ps->particles[cand[q]], vi, cd2, l0, e);
            }
            for (int d = 0; d < D; d++) after[(size_t)i * D + d] = vi[d];
        }
        free(cand);
    }
    free(outside);
    bins_free(&b);
#endif
    /* operators.c:435-443 */
    for (int i = 0; i < n; i++) {
        Particle *p = &ps->particles[i];
        if (p->type != FLUID_PARTICLE) continue;
        for (int d = 0; d < D; d++) {
            double dv = after[(size_t)i * D + d] - p->vel[d];
            p->pos[d] += dt * dv;
            p->vel[d] = after[(size_t)i * D + d];
        }
    }
    free(after);
}

/* clamp_pressure (operators.c:188-199) */
void orc_clamp(ParticleSystem *ps)
{
    for (int i = 0; i < ps->num; i++) {
        Particle *p = &ps->particles[i];
        if (p->type == FLUID_PARTICLE && p->pressure < 0.0) p->pressure = 0.0;
    }
}

/* calc_pressure_gradient (operators.c:102-181) then the corrector
 * (simulation.c:100-108) */
void orc_gradient_correct(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl)
{
    const int n = ps->num;
    const double re = c->influence_radius_lap, l0 = c->particle_distance;
    const double grad_coeff = (double)D / ps->n0;
    const double rho = c->density, dt = c->dt;
    const int mode = c->cmps_gradient;
    double *pmin = NULL;
    if (mode != 2) {
        pmin = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        if (!pmin) return;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            double m = ps->particles[i].pressure;
            for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
                double pj = ps->particles[nl->idx[k]].pressure;
                if (pj < m) m = pj;
            }
            pmin[i] = m;
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        const double pi_min = pmin ? pmin[i] : 0.0;
        double grad[D];
        for (int d = 0; d < D; d++) grad[d] = 0.0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const int j = nl->idx[k];
            const Particle *pj = &ps->particles[j];
            double dr[D];
            double r2 = dist2(pi, pj, dr);
            if (r2 < l0 * l0 * 1.0e-12) continue;
            double w = orc_weight(sqrt(r2), re, l0);
            double dp;
            if (mode == 1) dp = pi->pressure + pj->pressure - pi_min - pmin[j];
            else if (mode == 2) dp = pi->pressure + pj->pressure;
            else dp = pj->pressure - pi_min;
            for (int d = 0; d < D; d++) grad[d] += dp / r2 * dr[d] * w;
        }
        for (int d = 0; d < D; d++) pi->acc[d] = -grad_coeff * grad[d] / rho;
    }
    free(pmin);
    for (int i = 0; i < n; i++) {
        Particle *p = &ps->particles[i];
        if (p->type != FLUID_PARTICLE) continue;
        for (int d = 0; d < D; d++) {
            double du = dt * p->acc[d];
            p->vel[d] += du;
            p->pos[d] += dt * du;
        }
    }
}

/* apply_wall_boundary (boundary.c:12-26), remove_out_of_bounds (:33-61) */
void orc_boundary(ParticleSystem *ps, const SimConfig *c)
{
    const double margin = c->particle_distance * (c->wall_layers + 1);
    for (int i = 0; i < ps->num; i++) {
        Particle *p = &ps->particles[i];
        if (p->type == WALL_PARTICLE || p->type == GHOST_PARTICLE)
            for (int d = 0; d < D; d++) p->vel[d] = p->acc[d] = 0.0;
    }
    for (int i = 0; i < ps->num; i++) {
        Particle *p = &ps->particles[i];
        if (p->type != FLUID_PARTICLE) continue;
        int out = 0;
        for (int d = 0; d < D; d++)
            if (p->pos[d] < c->domain_min[d] - margin || p->pos[d] > c->domain_max[d] + margin) out = 1;
        for (int d = 0; d < D; d++)
            if (isnan(p->pos[d])) out = 1;
        if (!out) continue;
        p->type = GHOST_PARTICLE;
        p->pressure = 0.0;
        p->n = 0.0;
        for (int d = 0; d < D; d++) p->vel[d] = p->acc[d] = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* linear algebra                                                      */
/* ------------------------------------------------------------------ */

/* pressure_solver.c:79-84: serial sum; dot_order 1 sums backwards */
static double dot(const double *a, const double *b, int n, int order)
{
    double s = 0.0;
    if (order == 0)
        for (int i = 0; i < n; i++) s += a[i] * b[i];
    else
        for (int i = n - 1; i >= 0; i--) s += a[i] * b[i];
    return s;
}

/* csr_mat_vec (pressure_solver.c:57-68): row sums in column order */
static void spmv(int n, const long *rp, const int *col, const double *val,
                 const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (long p = rp[i]; p < rp[i + 1]; p++) s += val[p] * x[col[p]];
        y[i] = s;
    }
}

typedef struct {
    int n;
    const long *rp; const int *col; const double *val; const double *diag;
    /* factor */
    long *lp; int *lc; double *lv; double *ld;   /* ld: 1/sqrt(d) (sparse) or sqrt(d) (dense) */
    long *up; int *uc; long *ui;                 /* transpose of L: row j lists i>j ascending, ui = index into lv */
    int dense;
} icf;

/* ic_factorize_sparse (pressure_solver.c:156-232): right-looking IC(0),
 * restated row by row -- entry (i,k) = (A_ik - sum_j L_ij L_kj) * dinv_k with
 * the subtraction done term by term in ascending j; d_k = A_kk - L_kp^2 term
 * by term; d <= 0 falls back to A_kk (if > 0) or 1e-30. */
static void ic_factor_sparse(icf *f)
{
    const int n = f->n;
    for (int i = 0; i < n; i++) {
        for (long p = f->lp[i]; p < f->lp[i + 1]; p++) {
            const int k = f->lc[p];
            double s = f->lv[p];     /* A_ik, equal bit for bit to A_ki */
            long a = f->lp[i], bk = f->lp[k];
            while (a < f->lp[i + 1] && bk < f->lp[k + 1]) {
                int ja = f->lc[a], jb = f->lc[bk];
                if (ja < jb) a++;
                else if (ja > jb) bk++;
                else { s -= f->lv[a] * f->lv[bk]; a++; bk++; }
            }
            f->lv[p] = s * f->ld[k];
        }
        double d = f->diag[i];
        for (long p = f->lp[i]; p < f->lp[i + 1]; p++) d -= f->lv[p] * f->lv[p];
        if (d <= 0.0) d = (f->diag[i] > 0.0) ? f->diag[i] : 1.0e-30;
        f->ld[i] = 1.0 / sqrt(d);
    }
}

/* ic_factorize (mps_fluid2D/src/pressure_solver.c:86-117), dense formulas on
 * the pattern |A_ij| > 0: sum of squares first, L_kk = sqrt(d), entries
 * (A_ik - sum) * (1/L_kk) only where |A_ik| >= 1e-30.  Written up-looking
 * by rows; the values are the same because every entry reads only earlier
 * columns. */
static void ic_factor_dense(icf *f)
{
    const int n = f->n;
    for (int i = 0; i < n; i++) {
        for (long p = f->lp[i]; p < f->lp[i + 1]; p++) {
            const int k = f->lc[p];
            const double aik = f->lv[p];
            if (fabs(aik) < 1.0e-30) { f->lv[p] = 0.0; continue; }
            double s = 0.0;
            long a = f->lp[i], bk = f->lp[k];
            while (a < p && bk < f->lp[k + 1]) {
                int ja = f->lc[a], jb = f->lc[bk];
                if (ja < jb) a++;
                else if (ja > jb) bk++;
                else { s += f->lv[a] * f->lv[bk]; a++; bk++; }
            }
            f->lv[p] = (aik - s) * (1.0 / f->ld[k]);
        }
        double s = 0.0;
        for (long p = f->lp[i]; p < f->lp[i + 1]; p++) s += f->lv[p] * f->lv[p];
        double d = f->diag[i] - s;
        if (d <= 0.0) d = f->diag[i] > 0.0 ? f->diag[i] : 1.0e-30;
        f->ld[i] = sqrt(d);
    }
}

/* strict lower part of A (pattern |A| > 0 for the dense form) + transpose */
static int ic_setup(icf *f, int n, const long *rp, const int *col, const double *val,
                    const double *diag, int dense)
{
    memset(f, 0, sizeof *f);
    f->n = n; f->rp = rp; f->col = col; f->val = val; f->diag = diag; f->dense = dense;
    f->lp = (long *)calloc((size_t)n + 1, sizeof(long));
    f->up = (long *)calloc((size_t)n + 1, sizeof(long));
    f->ld = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!f->lp || !f->up || !f->ld) return -1;
    for (int i = 0; i < n; i++) {
        long c = 0;
        for (long p = rp[i]; p < rp[i + 1]; p++)
            if (col[p] < i && (!dense || fabs(val[p]) > 0.0)) { c++; f->up[col[p] + 1]++; }
        f->lp[i + 1] = f->lp[i] + c;
    }
    long nl = f->lp[n];
    for (int j = 0; j < n; j++) f->up[j + 1] += f->up[j];
    f->lc = (int *)malloc(sizeof(int) * (size_t)(nl + 1));
    f->lv = (double *)malloc(sizeof(double) * (size_t)(nl + 1));
    f->uc = (int *)malloc(sizeof(int) * (size_t)(nl + 1));
    f->ui = (long *)malloc(sizeof(long) * (size_t)(nl + 1));
    long *fill = (long *)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
    if (!f->lc || !f->lv || !f->uc || !f->ui || !fill) { free(fill); return -1; }
    memcpy(fill, f->up, sizeof(long) * (size_t)n);
    for (int i = 0; i < n; i++) {
        long q = f->lp[i];
        for (long p = rp[i]; p < rp[i + 1]; p++)
            if (col[p] < i && (!dense || fabs(val[p]) > 0.0)) {
                f->lc[q] = col[p];
                f->lv[q] = val[p];
                long t = fill[col[p]]++;
                f->uc[t] = i;     /* rows i ascending because i is the outer loop */
                f->ui[t] = q;
                q++;
            }
    }
    free(fill);
    if (dense) ic_factor_dense(f); else ic_factor_sparse(f);
    return 0;
}

static void ic_free(icf *f)
{
    free(f->lp); free(f->lc); free(f->lv); free(f->ld);
    free(f->up); free(f->uc); free(f->ui);
}

/* z = (L L^T)^-1 r.  Sparse form: forward_solve_sparse (pressure_solver.c:
 * 235-243) and backward_solve_sparse (:246-255), whose column scatter is
 * restated as a gather over L^T visiting rows i in descending order.  Dense
 * form: forward_solve / backward_solve (mps_fluid2D/src/pressure_solver.c:
 * 120-141), sums first then divide by L_ii. */
static void ic_apply(const icf *f, const double *r, double *tmp, double *z)
{
    const int n = f->n;
    if (!f->dense) {
        for (int i = 0; i < n; i++) {
            double s = r[i];
            for (long p = f->lp[i]; p < f->lp[i + 1]; p++) s -= f->lv[p] * tmp[f->lc[p]];
            tmp[i] = s * f->ld[i];
        }
        for (int j = n - 1; j >= 0; j--) {
            double s = tmp[j];
            for (long t = f->up[j + 1] - 1; t >= f->up[j]; t--) s -= f->lv[f->ui[t]] * z[f->uc[t]];
            z[j] = s * f->ld[j];
        }
    } else {
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (long p = f->lp[i]; p < f->lp[i + 1]; p++) s += f->lv[p] * tmp[f->lc[p]];
            tmp[i] = (r[i] - s) / f->ld[i];
        }
        for (int j = n - 1; j >= 0; j--) {
            double s = 0.0;
            for (long t = f->up[j]; t < f->up[j + 1]; t++) s += f->lv[f->ui[t]] * z[f->uc[t]];
            z[j] = (tmp[j] - s) / f->ld[j];
        }
    }
}

typedef struct { int iters; double res; double bnorm; int breakdown; } solve_info;

/* solve_cg_sparse (pressure_solver.c:90-131) / solve_iccg_sparse (:261-322) */
static int solve_impl(int n, const long *rp, const int *col, const double *val, const double *diag,
                      const double *b, double *x, int solver, int dense_ic, int max_iter,
                      double tol, const orc_opts *o, icf *keep, solve_info *info)
{
    const int order = o ? o->dot_order : 0;
    const size_t sz = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double *r = (double *)malloc(sz), *p = (double *)malloc(sz), *ap = (double *)malloc(sz);
    double *z = (double *)malloc(sz), *tmp = (double *)malloc(sz);
    if (!r || !p || !ap || !z || !tmp) { free(r); free(p); free(ap); free(z); free(tmp); return -1; }
    const double bnorm = sqrt(dot(b, b, n, order));
    const double stop = (o && o->tol_mode == 1) ? tol * bnorm : tol;
    icf f;
    int use_ic = (solver == 1) && ic_setup(&f, n, rp, col, val, diag, dense_ic) == 0;
    memset(x, 0, sz);
    memcpy(r, b, sz);
    int it = 0, brk = 0;
    double rr;
    if (!use_ic) {
        memcpy(p, r, sz);
        rr = dot(r, r, n, order);
        for (it = 0; it < max_iter; it++) {
            if (sqrt(rr) < stop) break;
            spmv(n, rp, col, val, p, ap);
            double pap = dot(p, ap, n, order);
            if (fabs(pap) < 1.0e-30) { brk = 1; break; }
            double alpha = rr / pap;
            for (int i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * ap[i];
            }
            double rr_new = dot(r, r, n, order);
            if (rr < 1.0e-60) break;
            double beta = rr_new / rr;
            for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
            rr = rr_new;
        }
        rr = dot(r, r, n, order);
    } else {
        ic_apply(&f, r, tmp, z);
        memcpy(p, z, sz);
        double rz = dot(r, z, n, order);
        for (it = 0; it < max_iter; it++) {
            if (sqrt(dot(r, r, n, order)) < stop) break;
            spmv(n, rp, col, val, p, ap);
            double pap = dot(p, ap, n, order);
            if (fabs(pap) < 1.0e-30) { brk = 1; break; }
            double alpha = rz / pap;
            for (int i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * ap[i];
            }
            ic_apply(&f, r, tmp, z);
            double rz_new = dot(r, z, n, order);
            if (fabs(rz) < 1.0e-60) break;
            double beta = rz_new / rz;
            for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            rz = rz_new;
        }
        rr = dot(r, r, n, order);
        if (keep) *keep = f; else ic_free(&f);
    }
    if (info) { info->iters = it; info->res = sqrt(rr); info->bnorm = bnorm; info->breakdown = brk; }
    free(r); free(p); free(ap); free(z); free(tmp);
    return it;
}

int orc_solve_csr(int n, const long *row_ptr, const int *col, const double *val,
                  const double *b, double *x, int solver, int dense_ic,
                  int max_iter, double tol, const orc_opts *o, double *res_out)
{
    double *diag = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!diag) return -1;
    for (int i = 0; i < n; i++)
        for (long p = row_ptr[i]; p < row_ptr[i + 1]; p++)
            if (col[p] == i) diag[i] = val[p];
    solve_info info;
    int it = solve_impl(n, row_ptr, col, val, diag, b, x, solver, dense_ic, max_iter, tol, o, NULL, &info);
    if (res_out) *res_out = info.res;
    free(diag);
    return it;
}

/* ------------------------------------------------------------------ */
/* pressure Poisson equation                                           */
/* ------------------------------------------------------------------ */

static void ppe_release(orc_ppe *q)
{
    free(q->eq_idx); free(q->row_ptr); free(q->col); free(q->val);
    free(q->diag); free(q->b); free(q->x);
    free(q->l_ptr); free(q->l_col); free(q->l_val); free(q->l_diag);
    memset(q, 0, sizeof *q);
}

/* source_term_standard (pressure_solver.c:328-340) */
static double source_standard(const ParticleSystem *ps, int i, const SimConfig *c, const double *vel_div)
{
    const double n0 = ps->n0, dt = c->dt, dt2 = dt * dt;
    if (c->ppe_type == 1)
        return -c->density / dt * vel_div[i]
               + c->gamma_ppe * (c->density / dt2) * (ps->particles[i].n - n0) / n0;
    return c->relaxation_coeff * (c->density / dt2) * (ps->particles[i].n - n0) / n0;
}

/* source_term_hs (pressure_solver.c:342-362) */
static double source_hs(const ParticleSystem *ps, int i, const SimConfig *c, const orc_nlist *nl)
{
    const double re = c->influence_radius_lap;
    const Particle *pi = &ps->particles[i];
    double sum = 0.0;
    for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
        const Particle *pj = &ps->particles[nl->idx[k]];
        if (pj->type == DUMMY_PARTICLE) continue;
        double dr[D];
        double r2 = dist2(pi, pj, dr);
        if (r2 < 1.0e-24) continue;
        double r = sqrt(r2), dv = 0.0;
        for (int d = 0; d < D; d++) dv += dr[d] * (pj->vel[d] - pi->vel[d]);
        sum += re / (r * r * r) * dv;
    }
    return -(c->density / ps->n0 / c->dt) * sum;
}

/* velocity divergence of the weakly-compressible PPE (pressure_solver.c:381-409) */
static void velocity_divergence(const ParticleSystem *ps, const SimConfig *c,
                                const orc_nlist *nl, double *div)
{
    const double re = c->influence_radius_lap, l0 = c->particle_distance;
    for (int i = 0; i < ps->num; i++) {
        div[i] = 0.0;
        const Particle *pi = &ps->particles[i];
        if (pi->type != FLUID_PARTICLE) continue;
        double s = 0.0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const Particle *pj = &ps->particles[nl->idx[k]];
            if (pj->type == DUMMY_PARTICLE) continue;
            double dr[D];
            double r2 = dist2(pi, pj, dr);
            if (r2 < 1.0e-24) continue;
            double dv = 0.0;
            for (int d = 0; d < D; d++) dv += (pj->vel[d] - pi->vel[d]) * dr[d];
            s += dv / r2 * orc_weight(sqrt(r2), re, l0);
        }
        div[i] = (double)D / ps->n0 * s;
    }
}

static int cmp_col_val(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* solve_pressure (pressure_solver.c:368-570) */
int orc_solve_pressure(orc_sim *s, ParticleSystem *ps, const SimConfig *c)
{
    const orc_nlist *nl = &s->nl;
    orc_ppe *q = &s->ppe;
    ppe_release(q);
    const int n = ps->num;
    const double re = c->influence_radius_lap, l0 = c->particle_distance;
    const double coeff = 2.0 * D / (ps->n0 * ps->lambda);

    double *vel_div = NULL;
    if (c->ppe_type == 1 && c->hs_mode == 0) {
        vel_div = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
        if (!vel_div) return -1;
        velocity_divergence(ps, c, nl, vel_div);
    }
    /* unknowns: interior fluid and every wall particle, in index order (:411-422) */
    q->eq_idx = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int n_eq = 0;
    for (int i = 0; i < n; i++) {
        const Particle *p = &ps->particles[i];
        int unknown = (p->type == FLUID_PARTICLE && !p->on_surface) || p->type == WALL_PARTICLE;
        q->eq_idx[i] = unknown ? n_eq++ : -1;
    }
    q->n_eq = n_eq;
    if (n_eq == 0) {
        for (int i = 0; i < n; i++) ps->particles[i].pressure = 0.0;
        free(vel_div);
        return 0;
    }
    /* row lengths and pointers (:437-463) */
    q->row_ptr = (long *)calloc((size_t)n_eq + 1, sizeof(long));
    for (int i = 0; i < n; i++) {
        if (q->eq_idx[i] < 0) continue;
        long cnt = 1;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            int j = nl->idx[k];
            if (ps->particles[j].type != DUMMY_PARTICLE && q->eq_idx[j] >= 0) cnt++;
        }
        q->row_ptr[q->eq_idx[i] + 1] = cnt;
    }
    for (int e = 0; e < n_eq; e++) q->row_ptr[e + 1] += q->row_ptr[e];
    q->nnz = q->row_ptr[n_eq];
    q->col = (int *)malloc(sizeof(int) * (size_t)q->nnz);
    q->val = (double *)malloc(sizeof(double) * (size_t)q->nnz);
    q->diag = (double *)calloc((size_t)n_eq, sizeof(double));
    q->b = (double *)calloc((size_t)n_eq, sizeof(double));
    q->x = (double *)calloc((size_t)n_eq, sizeof(double));
    /* fill (:490-531) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const int e = q->eq_idx[i];
        if (e < 0) continue;
        const Particle *pi = &ps->particles[i];
        long pos = q->row_ptr[e];
        double sum_w = 0.0;
        for (long k = nl->start[i]; k < nl->start[i + 1]; k++) {
            const int j = nl->idx[k];
            const Particle *pj = &ps->particles[j];
            if (pj->type == DUMMY_PARTICLE) continue;
            double dr[D];
            double w = orc_weight(sqrt(dist2(pi, pj, dr)), re, l0);
            sum_w += w;
            if (q->eq_idx[j] >= 0) {
                q->col[pos] = q->eq_idx[j];
                q->val[pos] = -coeff * w;
                pos++;
            }
        }
        double dv = (c->ppe_type == 1) ? c->c_ppe * coeff * sum_w : coeff * sum_w;
        q->col[pos] = e;
        q->val[pos] = dv;
        q->diag[e] = dv;
        if (pi->type == WALL_PARTICLE) q->b[e] = 0.0;
        else if (c->hs_mode == 1) q->b[e] = source_hs(ps, i, c, nl);
        else q->b[e] = source_standard(ps, i, c, vel_div);
    }
    /* rows sorted by column (:536-551); columns within a row are distinct */
    {
        int maxlen = 0;
        for (int e = 0; e < n_eq; e++) {
            int len = (int)(q->row_ptr[e + 1] - q->row_ptr[e]);
            if (len > maxlen) maxlen = len;
        }
#pragma omp parallel
        {
            struct cv { int c; int pad; double v; } *tmp =
                (struct cv *)malloc(sizeof(struct cv) * (size_t)(maxlen > 0 ? maxlen : 1));
#pragma omp for schedule(static)
            for (int e = 0; e < n_eq; e++) {
                long lo = q->row_ptr[e];
                int len = (int)(q->row_ptr[e + 1] - lo);
                for (int t = 0; t < len; t++) { tmp[t].c = q->col[lo + t]; tmp[t].v = q->val[lo + t]; }
                qsort(tmp, (size_t)len, sizeof(struct cv), cmp_col_val);
                for (int t = 0; t < len; t++) { q->col[lo + t] = tmp[t].c; q->val[lo + t] = tmp[t].v; }
            }
            free(tmp);
        }
    }
    free(vel_div);
    /* solve (:553-561) */
    solve_info info;
    icf f;
    int rc = solve_impl(n_eq, q->row_ptr, q->col, q->val, q->diag, q->b, q->x,
                        c->solver_type, D == 2, c->cg_max_iter, c->cg_tolerance,
                        &s->opts, c->solver_type == 1 ? &f : NULL, &info);
    if (rc < 0) return -1;
    q->iters = info.iters;
    q->res = info.res;
    q->bnorm = info.bnorm;
    q->breakdown = info.breakdown;
    if (c->solver_type == 1) {
        q->l_ptr = f.lp; q->l_col = f.lc; q->l_val = f.lv; q->l_diag = f.ld;
        free(f.up); free(f.uc); free(f.ui);
    }
    /* scatter (:563-566) */
    for (int i = 0; i < n; i++)
        ps->particles[i].pressure = (q->eq_idx[i] >= 0) ? q->x[q->eq_idx[i]] : 0.0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* driver                                                              */
/* ------------------------------------------------------------------ */

orc_sim *orc_sim_create(void)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof *s);
    if (s) s->overflow = -1;
    return s;
}

void orc_sim_free(orc_sim *s)
{
    if (!s) return;
    orc_nlist_free(&s->nl);
    ppe_release(&s->ppe);
    free(s);
}

void orc_sim_set_opts(orc_sim *s, int tol_mode, int dot_order)
{
    s->opts.tol_mode = tol_mode;
    s->opts.dot_order = dot_order;
}

int orc_sim_nlist(const orc_sim *s, const long **start, const int **idx)
{
    *start = s->nl.start;
    *idx = s->nl.idx;
    return s->nl.n;
}

const orc_ppe *orc_sim_ppe(const orc_sim *s) { return &s->ppe; }

int orc_sim_begin(orc_sim *s, ParticleSystem *ps, const SimConfig *c)
{
    int ov = orc_search(ps, c, search_radius(c), &s->nl);
    s->overflow = ov > 0 ? ov - 1 : -1;
    return ov;
}

/* simulation_step (simulation.c:19-118; 2-D mps_fluid2D/src/simulation.c:25-118) */
int orc_step(orc_sim *s, ParticleSystem *ps, const SimConfig *c)
{
    orc_visc_predict(ps, c, &s->nl);
#if MPS_DIM == 3
    if (c->collision_distance_ratio > 0.0)
#endif
        orc_collision(ps, c, &s->nl);
    int ov = orc_search(ps, c, search_radius(c), &s->nl);
    if (ov != 0) { s->overflow = ov - 1; return ov; }
    orc_density(ps, c, &s->nl);
    if (c->surface_detection_method == 1) {
        orc_neighbor_count(ps, c, &s->nl);
        orc_surface_by_count(ps, c->surface_count_threshold);
    } else {
        orc_surface(ps, c->surface_threshold);
    }
    if (orc_solve_pressure(s, ps, c) != 0) return -1;
    if (c->clamp_negative_pressure) orc_clamp(ps);
    orc_gradient_correct(ps, c, &s->nl);
    orc_boundary(ps, c);
    return 0;
}

/* ------------------------------------------------------------------ */
/* initial parameters (particle.c:69-221)                              */
/* ------------------------------------------------------------------ */

/* The reference sums, for each fluid i, w(r_ij) over every j != i in
 * ascending j where the weight is positive (particle.c:104-124) -- i.e. over
 * the j closer than the radius.  Those j are collected from a grid whose
 * cells are slightly larger than the radius (so no pair within it can fall
 * outside the 3x3(x3) block) and summed in ascending j: the same terms in
 * the same order. */
void orc_init_params(ParticleSystem *ps, const SimConfig *c)
{
    const int n = ps->num;
    const double re_n = c->influence_radius_n, re_lap = c->influence_radius_lap;
    const double re_st = c->influence_radius_st, l0 = c->particle_distance;
    const double reach = fmax(re_n, re_lap);
    double lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
    int any = 0;
    for (int i = 0; i < n; i++) {
        const double *x = ps->particles[i].pos;
        int finite = 1;
        for (int d = 0; d < D; d++) if (!isfinite(x[d])) finite = 0;
        if (!finite) continue;
        for (int d = 0; d < D; d++) {
            if (!any || x[d] < lo[d]) lo[d] = x[d];
            if (!any || x[d] > hi[d]) hi[d] = x[d];
        }
        any = 1;
    }
    orc_grid g;
    const double cell = reach * 1.0001;
    g.size = cell;
    g.inv = 1.0 / cell;
    g.total = 1;
    for (int d = 0; d < 3; d++) {
        if (d < D) {
            g.origin[d] = lo[d] - cell;
            g.nc[d] = (int)((hi[d] - lo[d]) / cell) + 3;
        } else {
            g.origin[d] = 0.0;
            g.nc[d] = 1;
        }
        g.total *= g.nc[d];
    }
    bins b;
    if (bins_build(&b, &g, ps, 1) != 0) return;

    double *n0_i = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *num_i = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *den_i = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    int *cnt_i = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    const double re_n2 = re_n * re_n;
#pragma omp parallel
    {
        int cap = 1024;
        int *hit = (int *)malloc(sizeof(int) * (size_t)cap);
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            const Particle *pi = &ps->particles[i];
            if (pi->type != FLUID_PARTICLE) continue;
            int cc[3];
            if (cell_coords(&g, pi->pos, cc) < 0) continue;   /* non-finite: every w is 0 or NaN */
            int clo[3], chi[3], m = 0;
            cell_range(&g, cc, clo, chi);
            for (int cz = clo[2]; cz <= chi[2]; cz++)
                for (int cy = clo[1]; cy <= chi[1]; cy++)
                    for (int cx = clo[0]; cx <= chi[0]; cx++) {
                        long cell_id = ((long)cz * g.nc[1] + cy) * g.nc[0] + cx;
                        for (long k = b.first[cell_id]; k < b.first[cell_id + 1]; k++) {
                            int j = b.member[k];
                            double dr[D];
                            if (j == i || !(dist2(pi, &ps->particles[j], dr) < reach * reach * 1.0001)) continue;
                            if (m == cap) { cap *= 2; hit = (int *)realloc(hit, sizeof(int) * (size_t)cap); }
                            hit[m++] = j;
                        }
                    }
            qsort(hit, (size_t)m, sizeof(int), cmp_int);
            double s_n = 0.0, s_num = 0.0, s_den = 0.0;
            int cnt = 0;
            for (int q = 0; q < m; q++) {
                double dr[D];
                double r2 = dist2(pi, &ps->particles[hit[q]], dr);
                double r = sqrt(r2);
                double wn = orc_weight(r, re_n, l0);
                if (wn > 0.0) s_n += wn;
                double wl = orc_weight(r, re_lap, l0);
                if (wl > 0.0) {
                    s_num += r2 * wl;
                    s_den += wl;
                }
                if (r2 < re_n2) cnt++;
            }
            n0_i[i] = s_n;
            num_i[i] = s_num;
            den_i[i] = s_den;
            cnt_i[i] = cnt;
        }
        free(hit);
    }
    bins_free(&b);

    double best = 0.0, best_num = 0.0, best_den = 0.0;
    int found = 0, max_count = 0;
    for (int i = 0; i < n; i++) {
        if (ps->particles[i].type != FLUID_PARTICLE) continue;
        if (n0_i[i] > best) {
            best = n0_i[i];
            best_num = num_i[i];
            best_den = den_i[i];
            found = 1;
        }
        if (cnt_i[i] > max_count) max_count = cnt_i[i];
    }
    free(n0_i); free(num_i); free(den_i); free(cnt_i);
    if (!found) {
        fprintf(stderr, "Error: no fluid particle found for initial params\n");
        ps->n0 = 1.0;
        ps->lambda = 1.0;
        return;
    }
    ps->n0 = best;
    if (c->surface_detection_method == 1) ps->n0_count = max_count;
    if (c->use_analytical_lambda)
        ps->lambda = re_lap * re_lap * (double)D * (double)(D - 1) / ((double)(D + 1) * (double)(D + 2));
    else
        ps->lambda = (best_den > 1.0e-10) ? best_num / best_den : 1.0;
    if (c->surface_tension_enabled) {
        /* C_LL from the lattice sum across a flat interface (particle.c:188-212;
         * 2-D: mps_fluid2D/src/particle.c, sigma*l0/sum) */
        int nmax = (int)ceil(re_st / l0);
        double sum = 0.0;
        for (int ax = 1; ax <= nmax; ax++)
            for (int bx = -(nmax - 1); bx <= 0; bx++)
                for (int by = -nmax; by <= nmax; by++)
#if MPS_DIM == 3
                    for (int bz = -nmax; bz <= nmax; bz++)
#endif
                    {
                        double rx = (double)(ax - bx) * l0;
                        double ry = (double)by * l0;
#if MPS_DIM == 3
                        double rz = (double)bz * l0;
                        double rab = sqrt(rx * rx + ry * ry + rz * rz);
#else
                        double rab = sqrt(rx * rx + ry * ry);
#endif
                        if (rab > 0.0 && rab < re_st)
                            sum += (1.0 / 3.0) * (rab - 1.5 * l0 + 0.5 * re_st) * (rab - re_st) * (rab - re_st);
                    }
        if (sum > 0.0) {
#if MPS_DIM == 3
            ps->C_LL = c->surface_tension_coeff * l0 * l0 / sum;
#else
            ps->C_LL = c->surface_tension_coeff * l0 / sum;
#endif
            double theta = c->wetting_angle_SL * M_PI / 180.0;
            ps->C_SL = 0.5 * (1.0 + cos(theta)) * ps->C_LL;
        }
    }
}

/* ------------------------------------------------------------------ */
/* particles.txt loader                                                */
/* ------------------------------------------------------------------ */

/* particle.c:38-58: capacity-bounded append */
static void add_particle(ParticleSystem *ps, const double pos[D], const double vel[D], int type)
{
    if (ps->num >= ps->capacity) {
        fprintf(stderr, "Error: particle capacity exceeded (%d)\n", ps->capacity);
        return;
    }
    Particle *p = &ps->particles[ps->num];
    for (int d = 0; d < D; d++) {
        p->pos[d] = pos[d];
        p->vel[d] = vel[d];
        p->acc[d] = 0.0;
    }
    p->pressure = 0.0;
    p->n = 0.0;
    p->type = type;
    p->on_surface = 0;
    ps->num++;
}

/* main.c:20-85 (2-D: mps_fluid2D/src/main.c:20-85, five columns), a line
 * at a time through a 512-byte fgets buffer: blanks and tabs skipped, '#',
 * newline, CR and empty lines ignored; the first line sscanf reads an int
 * from is the count; every row sscanf fills completely is appended, others
 * are reported and skipped; reading stops once the count is loaded. */
int orc_load_particles(const char *path, ParticleSystem *ps)
{
    FILE *fp = fopen(path, "r");
    if (!fp) {
        fprintf(stderr, "Error: cannot open particle file '%s'\n", path);
        return -1;
    }
    char line[512];
    int count = 0, have = 0;
    while (fgets(line, sizeof line, fp)) {
        char *s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '#' || *s == '\n' || *s == '\r' || *s == '\0') continue;
        if (sscanf(s, "%d", &count) == 1) {
            have = 1;
            break;
        }
    }
    if (!have || count <= 0) {
        fprintf(stderr, "Error: invalid particle count in '%s'\n", path);
        fclose(fp);
        return -1;
    }
    int loaded = 0, n_fluid = 0, n_wall = 0;
    while (fgets(line, sizeof line, fp) && loaded < count) {
        char *s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '#' || *s == '\n' || *s == '\r' || *s == '\0') continue;
        double pos[D], vel[D];
        int type;
#if D == 3
        if (sscanf(s, "%lf %lf %lf %lf %lf %lf %d", &pos[0], &pos[1], &pos[2],
                   &vel[0], &vel[1], &vel[2], &type) != 7) {
#else
        if (sscanf(s, "%lf %lf %lf %lf %d", &pos[0], &pos[1], &vel[0], &vel[1], &type) != 5) {
#endif
            fprintf(stderr, "Warning: skipping malformed line: %s", line);
            continue;
        }
        add_particle(ps, pos, vel, type);
        loaded++;
        if (type == FLUID_PARTICLE) n_fluid++;
        else if (type == WALL_PARTICLE) n_wall++;
    }
    fclose(fp);
    if (loaded != count)
        fprintf(stderr, "Warning: expected %d particles but loaded %d\n", count, loaded);
    printf("Loaded particles: %d fluid, %d wall, %d total\n", n_fluid, n_wall, ps->num);
    return 0;
}
