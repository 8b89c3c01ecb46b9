/*
 * nbx_oracle.c -- CPU restatement of the MichaelSt98/NNS Barnes-Hut path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (nns_amd/) never links it and
 * never falls back to it.
 *
 * Contents, each with the reference lines it restates:
 *   orc_bbox             bounding box + FP64 extent
 *                        (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:130-147, with the
 *                        min/max initialisation bug fixed; 3D/CUDA/CUDA_NBody/src/Kernels.cu:67-154)
 *   orc_keys_morton      Body::discretizePosition + Body::getKey
 *                        (3D/OpenMP/OpenMP_NBody/src/Body.cpp:13-32, include/Body.h:24,32)
 *   orc_lebesgue2hilbert Lebesgue2Hilbert (MolecularDynamics/BarnesHutParallel/
 *                        SpaceFillingCurves/src/Tree.cpp:70-104, maxlevel 21 from include/Tree.h:59)
 *   orc_sort_stable      ascending stable (key, index) order; the reference orders bodies by
 *                        key through std::map (Interaction.cpp:156-177)
 *   orc_tree_*           key-prefix octree (cells = key prefixes shared by >= 2 bodies),
 *                        FP64 bottom-up mass/COM, and the per-particle Barnes-Hut walk of
 *                        Interaction::treeInteraction (Interaction.cpp:58-98): a leaf is
 *                        always evaluated directly (distance 0 skipped, :28-30), a cell is
 *                        accepted when l/D < theta with l the cell side and D the distance
 *                        to its centre of mass, else its children are visited in slot order.
 *   orc_geo_*            the reference's own pointer octree restated: cubic Octant with
 *                        closed contains() (Octant.cpp:28-38), getSubOctant with '<='
 *                        tie-breaks (Octant.cpp:39-60), insert version 3 with the
 *                        incremental centre-of-mass update and push-down
 *                        (Tree.cpp:71-97,229-277), child visiting order UNW..LSE
 *                        (Interaction.cpp:73-96) and the SI-unit force of
 *                        singleInteraction (Interaction.cpp:17-55).
 *   orc_group_walk       work (child steps, frames) of the GPU's wavefront walk of the same
 *                        lists for groups of up to 64 bodies; no reference counterpart (the
 *                        reference walks one body at a time), pinned by the identities of a
 *                        one-body group (steps = interactions + frames + 1 at theta <= 0.5).
 *   orc_direct_accel     FP64 direct sum, the closed form a_i = sum_j m_j d_ij (|d_ij|^2+eps^2)^-3/2.
 *   orc_direct_accel_fast the same sum vectorised over sources (summation order differs).
 *   orc_star_forces      the reference's suns: bodies outside the tree, direct and
 *                        symmetric (Interaction.cpp:111-124), SI or Plummer G = 1 form
 *   orc_update_reference Interaction::updateBodies (Interaction.cpp:219-255)
 *   orc_update           symplectic Euler of updateKernel (Kernels.cu:581-600), written with
 *                        explicit fused multiply-adds so it is bit-reproducible.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_KEY_LEVELS 21
#define ORC_MASK_COORD 0x1FFFFFu /* 2^21 - 1, Body.h:32 */

int orc_version(void) { return 3; }

/* ------------------------------------------------------------------ bbox */
/* mm = {minX, maxX, minY, maxY, minZ, maxZ} as float (exact extremes of the
 * inputs); ext = {xMin, yMin, zMin, maxSpan} in FP64 (Interaction.cpp:143-147). */
void orc_bbox(int64_t n, const float *x, const float *y, const float *z, float *mm, double *ext)
{
    float a = x[0], b = x[0], c = y[0], d = y[0], e = z[0], f = z[0];
    for (int64_t i = 1; i < n; ++i) {
        a = x[i] < a ? x[i] : a; b = x[i] > b ? x[i] : b;
        c = y[i] < c ? y[i] : c; d = y[i] > d ? y[i] : d;
        e = z[i] < e ? z[i] : e; f = z[i] > f ? z[i] : f;
    }
    mm[0] = a; mm[1] = b; mm[2] = c; mm[3] = d; mm[4] = e; mm[5] = f;
    double xs = fabs((double)b - (double)a);
    double ys = fabs((double)d - (double)c);
    double zs = fabs((double)f - (double)e);
    double s = xs > ys ? xs : ys;
    s = s < zs ? zs : s;
    ext[0] = a; ext[1] = c; ext[2] = e; ext[3] = s;
}

/* ------------------------------------------------------------------ keys */
/* i_d = (uint)(MASK * (p_d - o_d) / maxSpan), evaluated left to right in FP64
 * and truncated (Body.cpp:14-16).  maxSpan == 0 (a single point) has no
 * defined value in the reference; here every coordinate then maps to 0. */
static inline uint64_t quantise(double p, double o, double s)
{
    if (!(s > 0.0)) return 0;
    double t = (double)ORC_MASK_COORD * (p - o);
    t = t / s;
    return (uint64_t)t;
}

static inline uint64_t interleave(uint64_t ix, uint64_t iy, uint64_t iz)
{
    /* key = 2^63 + sum_j 8^j (x_j + 2 y_j + 4 z_j) (Body.cpp:19-31) */
    uint64_t key = (uint64_t)1 << 63;
    ix &= ORC_MASK_COORD; iy &= ORC_MASK_COORD; iz &= ORC_MASK_COORD;
    for (int j = 0; j < ORC_KEY_LEVELS; ++j) {
        uint64_t digit = ((ix >> j) & 1u) + 2u * ((iy >> j) & 1u) + 4u * ((iz >> j) & 1u);
        key += digit << (3 * j);
    }
    return key;
}

uint64_t orc_morton_key(float x, float y, float z, const double *ext)
{
    return interleave(quantise(x, ext[0], ext[3]), quantise(y, ext[1], ext[3]),
                      quantise(z, ext[2], ext[3]));
}

void orc_keys_morton(int64_t n, const float *x, const float *y, const float *z,
                     const double *ext, uint64_t *keys)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) keys[i] = orc_morton_key(x[i], y[i], z[i], ext);
}

/* Hilbert tables: the state machine of Lebesgue2Hilbert (SFC/src/Tree.cpp:71-80).
 * Values must be identical for the keys to be the reference's keys. */
static const unsigned char ORC_DIR[12][8] = {
    {8, 10, 3, 3, 4, 5, 4, 5}, {2, 2, 11, 9, 4, 5, 4, 5}, {7, 6, 7, 6, 8, 10, 1, 1},
    {7, 6, 7, 6, 0, 0, 11, 9}, {0, 8, 1, 11, 6, 8, 6, 11}, {10, 0, 9, 1, 10, 7, 9, 7},
    {10, 4, 9, 4, 10, 2, 9, 3}, {5, 8, 5, 11, 2, 8, 3, 11}, {4, 9, 0, 0, 7, 9, 2, 2},
    {1, 1, 8, 5, 3, 3, 8, 6}, {11, 5, 0, 0, 11, 6, 2, 2}, {1, 1, 4, 10, 3, 3, 7, 10}};
static const unsigned char ORC_HIL[12][8] = {
    {0, 7, 3, 4, 1, 6, 2, 5}, {4, 3, 7, 0, 5, 2, 6, 1}, {6, 1, 5, 2, 7, 0, 4, 3},
    {2, 5, 1, 6, 3, 4, 0, 7}, {0, 1, 7, 6, 3, 2, 4, 5}, {6, 7, 1, 0, 5, 4, 2, 3},
    {2, 3, 5, 4, 1, 0, 6, 7}, {4, 5, 3, 2, 7, 6, 0, 1}, {0, 3, 1, 2, 7, 4, 6, 5},
    {2, 1, 3, 0, 5, 6, 4, 7}, {4, 7, 5, 6, 3, 0, 2, 1}, {6, 5, 7, 4, 1, 2, 0, 3}};

/* Lebesgue2Hilbert(lebesgue, level): lebesgue carries no placeholder bit; the
 * digits of levels deeper than `level` are emitted as 0 (SFC/src/Tree.cpp:82-104). */
uint64_t orc_lebesgue2hilbert(uint64_t lebesgue, int level)
{
    uint64_t h = 0;
    int dir = 0;
    for (int lvl = ORC_KEY_LEVELS; lvl > 0; --lvl) {
        int cell = (int)((lebesgue >> ((lvl - 1) * 3)) & 7u);
        h <<= 3;
        if (lvl > ORC_KEY_LEVELS - level) h += ORC_HIL[dir][cell];
        dir = ORC_DIR[dir][cell];
    }
    return h;
}

/* Hilbert key of a Morton key: placeholder bit removed, all 21 levels. */
void orc_keys_hilbert(int64_t n, const uint64_t *morton, uint64_t *out)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = orc_lebesgue2hilbert(morton[i] & (((uint64_t)1 << 63) - 1), ORC_KEY_LEVELS);
}

/* --------------------------------------------------------------- sorting */
/* Stable LSD radix sort, 8 bits per pass: ties keep ascending body index. */
void orc_sort_stable(int64_t n, const uint64_t *keys, int32_t *perm, uint64_t *sorted_keys)
{
    uint64_t *ka = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
    uint64_t *kb = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
    int32_t *va = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    int32_t *vb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) { ka[i] = keys[i]; va[i] = (int32_t)i; }
    for (int pass = 0; pass < 8; ++pass) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int sh = pass * 8;
        for (int64_t i = 0; i < n; ++i) cnt[((ka[i] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; ++i) {
            int64_t p = cnt[(ka[i] >> sh) & 255u]++;
            kb[p] = ka[i]; vb[p] = va[i];
        }
        uint64_t *tk = ka; ka = kb; kb = tk;
        int32_t *tv = va; va = vb; vb = tv;
    }
    for (int64_t i = 0; i < n; ++i) { perm[i] = va[i]; if (sorted_keys) sorted_keys[i] = ka[i]; }
    free(ka); free(kb); free(va); free(vb);
}

/* ------------------------------------------------------ key-prefix octree */
/* A cell at level L is a prefix of L key digits shared by at least two
 * bodies; its slot s holds the bodies whose digit L equals s (digit L is key
 * bits 62-3L .. 60-3L, so slot order is key order: slot = x + 2y + 4z with 1
 * for the upper half, Body.cpp:29).  A body is a leaf at the first level where
 * its prefix is unique.  Bodies with identical 63-bit keys end in a level-21
 * "bucket" cell whose slots take them in sorted order (at most 8).
 * Child encoding: >= 0 body index, -1 empty, <= -2 cell id (-2 - cid). */
typedef struct {
    int64_t n, ncell, cap;
    int32_t *level, *start, *count, *parent, *child; /* child: 8 per cell */
    double *mass, *mom;                              /* FP64 mass, sum m*x (3 per cell) */
    double *com;                                     /* FP64 centre of mass (3 per cell) */
    float *c32;                                      /* float x,y,z,m per cell */
    int32_t *perm;                                   /* sorted position -> body */
    uint64_t *skeys;
    float *x, *y, *z, *m;                            /* copies of the inputs */
    double ext[4];
    double lcell0;                                   /* side of the root cell */
    int err;
} orc_tree;

#define ORC_ERR_CAPACITY (-1)
#define ORC_ERR_BUCKET (-2)

static inline int digit_at(uint64_t key, int level) { return (int)((key >> (60 - 3 * level)) & 7u); }

static int64_t tree_new_cell(orc_tree *t, int level, int64_t lo, int64_t hi, int32_t parent)
{
    if (t->ncell >= t->cap) { t->err = ORC_ERR_CAPACITY; return -1; }
    int64_t c = t->ncell++;
    t->level[c] = level; t->start[c] = (int32_t)lo; t->count[c] = (int32_t)(hi - lo);
    t->parent[c] = parent;
    for (int s = 0; s < 8; ++s) t->child[8 * c + s] = -1;
    return c;
}

static void tree_summarise(orc_tree *t, int64_t c)
{
    /* FP64 sums over the children in slot order; bodies contribute m and m*x
     * computed exactly in FP64 from the float inputs. */
    double ms = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (int s = 0; s < 8; ++s) {
        int32_t ch = t->child[8 * c + s];
        if (ch == -1) continue;
        if (ch >= 0) {
            double mb = (double)t->m[ch];
            ms = ms + mb;
            mx = mx + mb * (double)t->x[ch];
            my = my + mb * (double)t->y[ch];
            mz = mz + mb * (double)t->z[ch];
        } else {
            int64_t cc = -2 - (int64_t)ch;
            ms = ms + t->mass[cc];
            mx = mx + t->mom[3 * cc + 0];
            my = my + t->mom[3 * cc + 1];
            mz = mz + t->mom[3 * cc + 2];
        }
    }
    t->mass[c] = ms;
    t->mom[3 * c + 0] = mx; t->mom[3 * c + 1] = my; t->mom[3 * c + 2] = mz;
    t->com[3 * c + 0] = mx / ms; t->com[3 * c + 1] = my / ms; t->com[3 * c + 2] = mz / ms;
    t->c32[4 * c + 0] = (float)t->com[3 * c + 0];
    t->c32[4 * c + 1] = (float)t->com[3 * c + 1];
    t->c32[4 * c + 2] = (float)t->com[3 * c + 2];
    t->c32[4 * c + 3] = (float)ms;
}

/* Builds the subtree of sorted range [lo, hi) at `level`; returns the child reference. */
static int32_t tree_build_rec(orc_tree *t, int level, int64_t lo, int64_t hi, int32_t parent)
{
    if (hi - lo == 1) return t->perm[lo];
    int64_t c = tree_new_cell(t, level, lo, hi, parent);
    if (c < 0) return -1;
    if (level == ORC_KEY_LEVELS) {
        if (hi - lo > 8) { t->err = ORC_ERR_BUCKET; return -1; }
        for (int64_t k = lo; k < hi; ++k) t->child[8 * c + (k - lo)] = t->perm[k];
    } else {
        int64_t a = lo;
        while (a < hi) {
            int d = digit_at(t->skeys[a], level);
            int64_t b = a + 1;
            while (b < hi && digit_at(t->skeys[b], level) == d) ++b;
            int32_t ref = tree_build_rec(t, level + 1, a, b, (int32_t)c);
            if (t->err) return -1;
            t->child[8 * c + d] = ref;
            a = b;
        }
    }
    tree_summarise(t, c);
    return (int32_t)(-2 - c);
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->level); free(t->start); free(t->count); free(t->parent); free(t->child);
    free(t->mass); free(t->mom); free(t->com); free(t->c32); free(t->perm); free(t->skeys);
    free(t->x); free(t->y); free(t->z); free(t->m);
    free(t);
}

/* Build from raw float inputs.  cap = maximum number of cells (0 -> 2n+12000,
 * the reference's numNodes - n headroom, BarnesHut.cu:18).  *err receives 0
 * or ORC_ERR_*. */
orc_tree *orc_tree_build(int64_t n, const float *x, const float *y, const float *z, const float *m,
                         int64_t cap, int *err)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->n = n;
    t->cap = cap > 0 ? cap : 2 * n + 12000;
    size_t cn = (size_t)t->cap;
    t->level = malloc(sizeof(int32_t) * cn); t->start = malloc(sizeof(int32_t) * cn);
    t->count = malloc(sizeof(int32_t) * cn); t->parent = malloc(sizeof(int32_t) * cn);
    t->child = malloc(sizeof(int32_t) * 8 * cn);
    t->mass = malloc(sizeof(double) * cn); t->mom = malloc(sizeof(double) * 3 * cn);
    t->com = malloc(sizeof(double) * 3 * cn); t->c32 = malloc(sizeof(float) * 4 * cn);
    t->perm = malloc(sizeof(int32_t) * (size_t)n); t->skeys = malloc(sizeof(uint64_t) * (size_t)n);
    t->x = malloc(sizeof(float) * (size_t)n); t->y = malloc(sizeof(float) * (size_t)n);
    t->z = malloc(sizeof(float) * (size_t)n); t->m = malloc(sizeof(float) * (size_t)n);
    memcpy(t->x, x, sizeof(float) * (size_t)n); memcpy(t->y, y, sizeof(float) * (size_t)n);
    memcpy(t->z, z, sizeof(float) * (size_t)n); memcpy(t->m, m, sizeof(float) * (size_t)n);
    float mm[6];
    orc_bbox(n, x, y, z, mm, t->ext);
    t->lcell0 = ldexp(t->ext[3] / (double)ORC_MASK_COORD, ORC_KEY_LEVELS);
    uint64_t *keys = malloc(sizeof(uint64_t) * (size_t)n);
    orc_keys_morton(n, x, y, z, t->ext, keys);
    orc_sort_stable(n, keys, t->perm, t->skeys);
    free(keys);
    if (n >= 2) {
        /* the root is always a cell (cell 0), even when all bodies differ in digit 0 */
        int64_t root = tree_new_cell(t, 0, 0, n, -1);
        (void)root;
        int64_t a = 0;
        while (a < n && !t->err) {
            int d = digit_at(t->skeys[a], 0);
            int64_t b = a + 1;
            while (b < n && digit_at(t->skeys[b], 0) == d) ++b;
            int32_t ref = tree_build_rec(t, 1, a, b, 0);
            t->child[d] = ref;
            a = b;
        }
        if (!t->err) tree_summarise(t, 0);
    }
    *err = t->err;
    return t;
}

int64_t orc_tree_ncells(const orc_tree *t) { return t->ncell; }
double orc_tree_lcell0(const orc_tree *t) { return t->lcell0; }

void orc_tree_export(const orc_tree *t, int32_t *level, int32_t *start, int32_t *count,
                     int32_t *parent, int32_t *child, double *mass, double *com, float *c32,
                     int32_t *perm, uint64_t *skeys, double *ext)
{
    size_t c = (size_t)t->ncell, n = (size_t)t->n;
    if (level) memcpy(level, t->level, 4 * c);
    if (start) memcpy(start, t->start, 4 * c);
    if (count) memcpy(count, t->count, 4 * c);
    if (parent) memcpy(parent, t->parent, 4 * c);
    if (child) memcpy(child, t->child, 32 * c);
    if (mass) memcpy(mass, t->mass, 8 * c);
    if (com) memcpy(com, t->com, 24 * c);
    if (c32) memcpy(c32, t->c32, 16 * c);
    if (perm) memcpy(perm, t->perm, 4 * n);
    if (skeys) memcpy(skeys, t->skeys, 8 * n);
    if (ext) memcpy(ext, t->ext, 32);
}

/* Opening threshold of the FP32 predicate used on the GPU: crit_L = (l_L/theta)^2
 * as a float, crit_0 rounded once from FP64 and scaled by exact powers of 4; the
 * GPU compares crit_L + eps^2 (one float addition) with |d|^2 + eps^2. */
float orc_crit0(double lcell0, double theta)
{
    double r = lcell0 / theta;
    return (float)(r * r);
}

/* Per-particle Barnes-Hut walk (Interaction::treeInteraction, Interaction.cpp:58-98)
 * over the key-prefix tree with the Plummer kernel m d (|d|^2+eps2)^-3/2 and G = 1.
 *   mac_mode 0: reference predicate in FP64, l/D < theta, cell data in FP64.
 *   mac_mode 1: the GPU's FP32 predicate
 *                   crit_L + eps2 < fma(dz,dz,fma(dy,dy,fma(dx,dx,eps2)))
 *               (l/D < theta with eps^2 added to both sides) on float data, cell
 *               data taken as the float values the GPU stores.
 * Accumulation is FP64 in both modes.  ninter[i] counts the accepted cells plus the
 * leaf bodies other than i itself (the per-particle interaction count). */
int orc_tree_accel(const orc_tree *t, double theta, double eps2, int mac_mode,
                   int64_t ntarget, const int32_t *targets, double *acc, int64_t *ninter)
{
    if (t->n < 2) {
        for (int64_t k = 0; k < ntarget; ++k) {
            acc[3 * k] = acc[3 * k + 1] = acc[3 * k + 2] = 0.0;
            if (ninter) ninter[k] = 0;
        }
        return 0;
    }
    const float crit0 = orc_crit0(t->lcell0, theta);
    const float eps2f = (float)eps2;
    float critf[ORC_KEY_LEVELS + 2];
    double lcell[ORC_KEY_LEVELS + 2];
    for (int L = 0; L <= ORC_KEY_LEVELS + 1; ++L) {
        critf[L] = ldexpf(crit0, -2 * L) + eps2f;
        lcell[L] = ldexp(t->lcell0, -L);
    }
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < ntarget; ++k) {
        int32_t i = targets ? targets[k] : (int32_t)k;
        const float pxf = t->x[i], pyf = t->y[i], pzf = t->z[i];
        const double px = pxf, py = pyf, pz = pzf;
        double ax = 0.0, ay = 0.0, az = 0.0;
        int64_t cnt = 0;
        int32_t stk_node[8 * (ORC_KEY_LEVELS + 2)];
        int32_t sp = 0;
        stk_node[sp++] = -2; /* root cell 0 */
        while (sp > 0) {
            int32_t ref = stk_node[--sp];
            double mj, qx, qy, qz;
            if (ref >= 0) {
                if (ref != i) cnt++;
                mj = t->m[ref]; qx = t->x[ref]; qy = t->y[ref]; qz = t->z[ref];
            } else {
                int64_t c = -2 - (int64_t)ref;
                int L = t->level[c];
                int accept;
                if (mac_mode == 0) {
                    double dx = t->com[3 * c] - px, dy = t->com[3 * c + 1] - py,
                           dz = t->com[3 * c + 2] - pz;
                    double D = sqrt(dx * dx + dy * dy + dz * dz);
                    accept = (lcell[L] / D) < theta;
                    mj = t->mass[c]; qx = t->com[3 * c]; qy = t->com[3 * c + 1]; qz = t->com[3 * c + 2];
                } else {
                    float dx = t->c32[4 * c] - pxf, dy = t->c32[4 * c + 1] - pyf,
                          dz = t->c32[4 * c + 2] - pzf;
                    float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                    accept = critf[L] < d2;
                    mj = t->c32[4 * c + 3]; qx = t->c32[4 * c]; qy = t->c32[4 * c + 1];
                    qz = t->c32[4 * c + 2];
                }
                if (!accept) {
                    for (int s = 7; s >= 0; --s) {
                        int32_t ch = t->child[8 * c + s];
                        if (ch != -1) stk_node[sp++] = ch;
                    }
                    continue;
                }
                cnt++;
            }
            double dx = qx - px, dy = qy - py, dz = qz - pz;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0) continue; /* Interaction.cpp:28-30 */
            double r2 = d2 + eps2;
            double inv = 1.0 / (r2 * sqrt(r2));
            ax += mj * dx * inv; ay += mj * dy * inv; az += mj * dz * inv;
        }
        acc[3 * k] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;
        if (ninter) ninter[k] = cnt;
    }
    return 0;
}

/* ------------------------------------------------ work of the group walk */
/* Bottom of the one-child chain that starts at cell c: the first cell on the way
 * down with two or more non-empty children (a cell always holds two or more
 * bodies, so a one-child cell's child is a cell). */
static int64_t orc_chain_bottom(const orc_tree *t, int64_t c, int *nch)
{
    for (;;) {
        int k = 0;
        int32_t last = -1;
        for (int s = 0; s < 8; ++s)
            if (t->child[8 * c + s] != -1) { ++k; last = t->child[8 * c + s]; }
        if (k == 1 && last <= -2) { c = -2 - (int64_t)last; continue; }
        *nch = k;
        return c;
    }
}

/* Work of the wavefront walk that evaluates orc_tree_accel's mac_mode 1 lists
 * for groups of up to 64 bodies at once (one lane per body).  The walk tests a
 * cell child for all lanes of its parent's frame and opens it as a frame of
 * its own for the lanes that reject it; one-child chains are described by
 * their bottom cell and tested with its level's threshold (the chain's cells
 * share mass and centre, and the threshold falls with the level, so each body
 * accepts the chain or opens its bottom exactly as the per-body walk does).
 * A frame therefore exists for every chain-bottom cell that at least one lane
 * of the group opens -- the union of the lanes' opened sets -- and the child
 * steps of a group are 1 (the root's test) plus the children of its frames.
 * order[k] is the body in group-order position k; group g holds positions
 * [first[g], first[g] + cnt[g]).  Outputs steps[g], frames[g].  Returns -2 for
 * a tree of fewer than two bodies (it has no cell), -1 when out of memory. */
int orc_group_walk(const orc_tree *t, double theta, double eps2, int64_t ngroups, const int32_t *order,
                   const int64_t *first, const int32_t *cnt, int64_t *steps, int64_t *frames)
{
    if (t->n < 2) return -2;
    const float crit0 = orc_crit0(t->lcell0, theta);
    const float eps2f = (float)eps2;
    float critf[ORC_KEY_LEVELS + 2];
    for (int L = 0; L <= ORC_KEY_LEVELS + 1; ++L) critf[L] = ldexpf(crit0, -2 * L) + eps2f;
    const int64_t nc = t->ncell;
    int bad = 0;
    #pragma omp parallel
    {
        uint8_t *mark = (uint8_t *)calloc((size_t)(nc ? nc : 1), 1);
        int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc ? nc : 1));
        if (!mark || !touched) {
            #pragma omp atomic write
            bad = 1;
        }
        #pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < ngroups; ++g) {
            if (!mark || !touched) continue;
            int64_t nt = 0, st = 1, fr = 0;
            for (int32_t l = 0; l < cnt[g]; ++l) {
                const int32_t i = order[first[g] + l];
                const float pxf = t->x[i], pyf = t->y[i], pzf = t->z[i];
                int32_t stk[8 * (ORC_KEY_LEVELS + 2)];
                int32_t sp = 0;
                stk[sp++] = 0; /* the root cell, as a child of a virtual parent */
                while (sp > 0) {
                    int nch;
                    const int64_t b = orc_chain_bottom(t, stk[--sp], &nch);
                    const float dx = t->c32[4 * b] - pxf, dy = t->c32[4 * b + 1] - pyf,
                                dz = t->c32[4 * b + 2] - pzf;
                    const float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                    if (critf[t->level[b]] < d2) continue; /* accepted */
                    if (!mark[b]) {
                        mark[b] = 1;
                        touched[nt++] = (int32_t)b;
                        st += nch;
                        fr += 1;
                    }
                    for (int s = 0; s < 8; ++s) {
                        const int32_t ch = t->child[8 * b + s];
                        if (ch <= -2) stk[sp++] = -2 - ch;
                    }
                }
            }
            for (int64_t k = 0; k < nt; ++k) mark[touched[k]] = 0;
            steps[g] = st;
            frames[g] = fr;
        }
        free(mark);
        free(touched);
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------ vector-Jacobian product */
/* VJP of the accelerations of orc_tree_accel with the interaction lists held
 * fixed: given u[i] = dL/da_i, returns gx[k] = dL/dx_k and gm[k] = dL/dm_k.
 * Per accepted node n of body i (d = X_n - x_i, r2 = |d|^2 + eps2, s = r2^-3/2):
 *   h = M_n s (u_i - 3 (u_i.d) d / r2);  dL/dx_i -= h;  dL/dX_n += h;  dL/dM_n += s (u_i.d).
 * A leaf passes dL/dX, dL/dM straight to its body.  A cell passes them to every
 * body k below it through X = sum m x / M, M = sum m:
 *   dL/dx_k += m_k G_n / M_n,   dL/dm_k += GM_n + G_n.(x_k - X_n) / M_n,
 * accumulated top-down over the ancestors.  Zero-distance pairs contribute
 * nothing, as in the forward walk.  Serial; FP64 throughout (cell centres and
 * masses in FP64 for the chain, in the mode's precision for the interaction). */
int orc_tree_vjp(const orc_tree *t, double theta, double eps2, int mac_mode, const double *u, double *gx,
                 double *gm)
{
    const int64_t n = t->n, nc = t->ncell;
    for (int64_t k = 0; k < n; ++k) { gx[3 * k] = gx[3 * k + 1] = gx[3 * k + 2] = 0.0; gm[k] = 0.0; }
    if (n < 2) return 0;
    double *G = calloc((size_t)(4 * nc), sizeof(double)); /* Gx, Gy, Gz, GM per cell */
    const float crit0 = orc_crit0(t->lcell0, theta);
    const float eps2f = (float)eps2;
    float critf[ORC_KEY_LEVELS + 2];
    double lcell[ORC_KEY_LEVELS + 2];
    for (int L = 0; L <= ORC_KEY_LEVELS + 1; ++L) {
        critf[L] = ldexpf(crit0, -2 * L) + eps2f;
        lcell[L] = ldexp(t->lcell0, -L);
    }
    for (int64_t i = 0; i < n; ++i) {
        const float pxf = t->x[i], pyf = t->y[i], pzf = t->z[i];
        const double px = pxf, py = pyf, pz = pzf;
        const double ux = u[3 * i], uy = u[3 * i + 1], uz = u[3 * i + 2];
        int32_t stk[8 * (ORC_KEY_LEVELS + 2)];
        int32_t sp = 0;
        stk[sp++] = -2;
        while (sp > 0) {
            int32_t ref = stk[--sp];
            double mj, qx, qy, qz;
            int64_t c = -1;
            if (ref >= 0) {
                mj = t->m[ref]; qx = t->x[ref]; qy = t->y[ref]; qz = t->z[ref];
            } else {
                c = -2 - (int64_t)ref;
                int L = t->level[c];
                int accept;
                if (mac_mode == 0) {
                    double dx = t->com[3 * c] - px, dy = t->com[3 * c + 1] - py, dz = t->com[3 * c + 2] - pz;
                    accept = (lcell[L] / sqrt(dx * dx + dy * dy + dz * dz)) < theta;
                    mj = t->mass[c]; qx = t->com[3 * c]; qy = t->com[3 * c + 1]; qz = t->com[3 * c + 2];
                } else {
                    float dx = t->c32[4 * c] - pxf, dy = t->c32[4 * c + 1] - pyf, dz = t->c32[4 * c + 2] - pzf;
                    accept = critf[L] < fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                    mj = t->c32[4 * c + 3]; qx = t->c32[4 * c]; qy = t->c32[4 * c + 1]; qz = t->c32[4 * c + 2];
                }
                if (!accept) {
                    for (int s = 7; s >= 0; --s) {
                        int32_t ch = t->child[8 * c + s];
                        if (ch != -1) stk[sp++] = ch;
                    }
                    continue;
                }
            }
            /* The body's own term vanishes identically; another body at the same
             * point still has a derivative (m eps^-3) when eps > 0. */
            if (ref == i) continue;
            double dx = qx - px, dy = qy - py, dz = qz - pz;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0 && eps2 == 0.0) continue;
            double r2 = d2 + eps2;
            double s = 1.0 / (r2 * sqrt(r2));
            double ud = ux * dx + uy * dy + uz * dz;
            double k3 = 3.0 * ud / r2;
            double hx = mj * s * (ux - k3 * dx), hy = mj * s * (uy - k3 * dy), hz = mj * s * (uz - k3 * dz);
            gx[3 * i] -= hx; gx[3 * i + 1] -= hy; gx[3 * i + 2] -= hz;
            if (ref >= 0) {
                gx[3 * ref] += hx; gx[3 * ref + 1] += hy; gx[3 * ref + 2] += hz;
                gm[ref] += s * ud;
            } else {
                G[4 * c] += hx; G[4 * c + 1] += hy; G[4 * c + 2] += hz; G[4 * c + 3] += s * ud;
            }
        }
    }
    /* top-down: A = sum over ancestors of G/M, B = sum of GM - G.X/M */
    double *A = malloc(sizeof(double) * 4 * (size_t)nc);
    int64_t *stk = malloc(sizeof(int64_t) * (size_t)(nc + 1));
    int64_t sp = 0;
    stk[sp++] = 0;
    while (sp > 0) {
        int64_t c = stk[--sp];
        const double M = t->mass[c];
        double a0 = G[4 * c] / M, a1 = G[4 * c + 1] / M, a2 = G[4 * c + 2] / M;
        double b = G[4 * c + 3] - (a0 * t->com[3 * c] + a1 * t->com[3 * c + 1] + a2 * t->com[3 * c + 2]);
        int32_t par = t->parent[c];
        if (par >= 0) { a0 += A[4 * par]; a1 += A[4 * par + 1]; a2 += A[4 * par + 2]; b += A[4 * par + 3]; }
        A[4 * c] = a0; A[4 * c + 1] = a1; A[4 * c + 2] = a2; A[4 * c + 3] = b;
        for (int s = 0; s < 8; ++s) {
            int32_t ch = t->child[8 * c + s];
            if (ch == -1) continue;
            if (ch >= 0) {
                const double mk = t->m[ch];
                gx[3 * ch] += mk * a0; gx[3 * ch + 1] += mk * a1; gx[3 * ch + 2] += mk * a2;
                gm[ch] += b + a0 * (double)t->x[ch] + a1 * (double)t->y[ch] + a2 * (double)t->z[ch];
            } else {
                stk[sp++] = -2 - (int64_t)ch;
            }
        }
    }
    free(A); free(stk); free(G);
    return 0;
}

/* --------------------------------------------------------- direct sum */
void orc_direct_accel(int64_t n, const float *x, const float *y, const float *z, const float *m,
                      double eps2, int64_t ntarget, const int32_t *targets, double *acc)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < ntarget; ++k) {
        int32_t i = targets ? targets[k] : (int32_t)k;
        const double px = x[i], py = y[i], pz = z[i];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double dx = (double)x[j] - px, dy = (double)y[j] - py, dz = (double)z[j] - pz;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0) continue;
            double r2 = d2 + eps2;
            double inv = (double)m[j] / (r2 * sqrt(r2));
            ax += dx * inv; ay += dy * inv; az += dz * inv;
        }
        acc[3 * k] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;
    }
}

/* Same closed form for large samples: sources converted once to FP64 arrays and
 * the inner loop vectorised (AVX2 when the host has it).  Summation order
 * differs from orc_direct_accel, so results agree to FP64 rounding only. */
#define ORC_DIRECT_LOOP                                                                 \
    for (int64_t k = 0; k < ntarget; ++k) {                                             \
        int32_t i = targets ? targets[k] : (int32_t)k;                                  \
        const double px = x[i], py = y[i], pz = z[i];                                   \
        double ax = 0.0, ay = 0.0, az = 0.0;                                            \
        _Pragma("omp simd reduction(+ : ax, ay, az)")                                  \
        for (int64_t j = 0; j < n; ++j) {                                               \
            double dx = X[j] - px, dy = Y[j] - py, dz = Z[j] - pz;                      \
            double d2 = dx * dx + dy * dy + dz * dz;                                    \
            const int far = d2 > 0.0;                                                   \
            double r2 = far ? d2 + eps2 : 1.0;                                          \
            double w = (far ? M[j] : 0.0) / (r2 * sqrt(r2));                            \
            ax += dx * w; ay += dy * w; az += dz * w;                                   \
        }                                                                               \
        acc[3 * k] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;                      \
    }

__attribute__((target("avx2,fma"))) static void direct_avx2(
    int64_t n, const float *x, const float *y, const float *z, const double *X, const double *Y,
    const double *Z, const double *M, double eps2, int64_t ntarget, const int32_t *targets, double *acc)
{
    #pragma omp parallel for schedule(dynamic, 4)
    ORC_DIRECT_LOOP
}

static void direct_generic(
    int64_t n, const float *x, const float *y, const float *z, const double *X, const double *Y,
    const double *Z, const double *M, double eps2, int64_t ntarget, const int32_t *targets, double *acc)
{
    #pragma omp parallel for schedule(dynamic, 4)
    ORC_DIRECT_LOOP
}

/* Returns 0, or -1 when the FP64 copies cannot be allocated (acc untouched). */
int orc_direct_accel_fast(int64_t n, const float *x, const float *y, const float *z, const float *m,
                          double eps2, int64_t ntarget, const int32_t *targets, double *acc)
{
    const size_t bytes = 8 * (size_t)(n > 0 ? n : 1);
    double *X = malloc(bytes), *Y = malloc(bytes), *Z = malloc(bytes), *M = malloc(bytes);
    if (!X || !Y || !Z || !M) {
        free(X); free(Y); free(Z); free(M);
        return -1;
    }
    for (int64_t j = 0; j < n; ++j) { X[j] = x[j]; Y[j] = y[j]; Z[j] = z[j]; M[j] = m[j]; }
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
        direct_avx2(n, x, y, z, X, Y, Z, M, eps2, ntarget, targets, acc);
    else
        direct_generic(n, x, y, z, X, Y, Z, M, eps2, ntarget, targets, acc);
    free(X); free(Y); free(Z); free(M);
    return 0;
}

/* ------------------------------------------------------------- update */
/* v += dt a; x += d dt v  (Kernels.cu:590-596), as fma(dt,a,v) and fma(d*dt,v,x). */
void orc_update(int64_t n, float *x, float *y, float *z, float *vx, float *vy, float *vz,
                const float *ax, const float *ay, const float *az, float dt, float d)
{
    const float ddt = d * dt;
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = fmaf(dt, ax[i], vx[i]); vy[i] = fmaf(dt, ay[i], vy[i]); vz[i] = fmaf(dt, az[i], vz[i]);
        x[i] = fmaf(ddt, vx[i], x[i]); y[i] = fmaf(ddt, vy[i], y[i]); z[i] = fmaf(ddt, vz[i], z[i]);
    }
}

/* ============================================= reference pointer octree */
/* Restatement of Tree/Octant (3D/OpenMP/OpenMP_NBody/src/Tree.cpp, Octant.cpp)
 * on a node pool.  Sub-octant numbering is the reference's: 0 UNW, 1 UNE, 2 USW,
 * 3 USE, 4 LNW, 5 LNE, 6 LSW, 7 LSE (Octant.cpp:39-60). */
typedef struct { double mass, x, y, z; } gbody;
typedef struct {
    double cx, cy, cz, len;
    gbody com;
    int32_t child[8];
    int32_t level;
} gnode;
typedef struct {
    gnode *nodes;
    int64_t nn, cap;
    int err;
} orc_geo;

static int32_t geo_new(orc_geo *g, double cx, double cy, double cz, double len, int level)
{
    if (g->nn >= g->cap) {
        int64_t nc = g->cap * 2 + 1024;
        gnode *p = (gnode *)realloc(g->nodes, sizeof(gnode) * (size_t)nc);
        if (!p) { g->err = ORC_ERR_CAPACITY; return -1; }
        g->nodes = p; g->cap = nc;
    }
    gnode *q = &g->nodes[g->nn];
    q->cx = cx; q->cy = cy; q->cz = cz; q->len = len;
    q->com.mass = 0.0; q->com.x = q->com.y = q->com.z = 0.0;
    for (int s = 0; s < 8; ++s) q->child[s] = -1;
    q->level = level;
    return (int32_t)g->nn++;
}

static int geo_external(const gnode *q)
{
    for (int s = 0; s < 8; ++s) if (q->child[s] != -1) return 0;
    return 1;
}

static int geo_suboctant(const gnode *q, const gbody *b)
{
    /* Octant::getSubOctant: lower if z <= cz, south if y <= cy, west if x <= cx */
    int s = 0;
    if (b->z <= q->cz) s += 4;
    if (b->y <= q->cy) s += 2;
    if (!(b->x <= q->cx)) s += 1;
    return s;
}

/* Tree::insert, version 3 (Tree.cpp:71-97, 229-277); depth guard added
 * (the reference recurses forever on coincident bodies). */
static void geo_insert(orc_geo *g, int32_t node, gbody body, int depth)
{
    if (g->err) return;
    if (depth > 200) { g->err = ORC_ERR_BUCKET; return; }
    gnode *q = &g->nodes[node];
    if (q->com.mass == 0.0) { q->com = body; return; }
    int ext = geo_external(q);
    gbody upd;
    if (!ext) {
        q->com.x = (body.x * body.mass + q->com.x * q->com.mass) / (body.mass + q->com.mass);
        q->com.y = (body.y * body.mass + q->com.y * q->com.mass) / (body.mass + q->com.mass);
        q->com.z = (body.z * body.mass + q->com.z * q->com.mass) / (body.mass + q->com.mass);
        q->com.mass += body.mass;
        upd = body;
    } else {
        upd = q->com;
    }
    int s = geo_suboctant(q, &upd);
    if (q->child[s] == -1) {
        double nl = q->len / 4.0;
        double cx = q->cx + ((s & 1) ? nl : -nl);
        double cy = q->cy + ((s & 2) ? -nl : nl);
        double cz = q->cz + ((s & 4) ? -nl : nl);
        int32_t c = geo_new(g, cx, cy, cz, q->len / 2.0, q->level + 1);
        if (c < 0) return;
        q = &g->nodes[node]; /* pool may have moved */
        q->child[s] = c;
    }
    geo_insert(g, g->nodes[node].child[s], upd, depth + 1);
    if (ext) geo_insert(g, node, body, depth + 1);
}

static int geo_contains(const gnode *q, double x, double y, double z)
{
    /* Octant::contains (Octant.cpp:28-38) */
    return x <= (q->cx + q->len / 2.0) && x >= (q->cx - q->len / 2.0) &&
           y <= (q->cy + q->len / 2.0) && y >= (q->cy - q->len / 2.0) &&
           z <= (q->cz + q->len / 2.0) && z >= (q->cz - q->len / 2.0);
}

orc_geo *orc_geo_build(int64_t n, const double *x, const double *y, const double *z,
                       const double *m, double cx, double cy, double cz, double len, int *err)
{
    orc_geo *g = (orc_geo *)calloc(1, sizeof(orc_geo));
    g->cap = 2 * n + 16;
    g->nodes = (gnode *)malloc(sizeof(gnode) * (size_t)g->cap);
    geo_new(g, cx, cy, cz, len, 0);
    for (int64_t i = 0; i < n && !g->err; ++i) {
        if (!geo_contains(&g->nodes[0], x[i], y[i], z[i])) continue; /* Interaction.cpp:182-185 */
        gbody b = {m[i], x[i], y[i], z[i]};
        geo_insert(g, 0, b, 0);
    }
    *err = g->err;
    return g;
}

void orc_geo_free(orc_geo *g)
{
    if (!g) return;
    free(g->nodes);
    free(g);
}

int64_t orc_geo_nnodes(const orc_geo *g) { return g->nn; }

/* Node dump in pool order: per node {level, is_external} and {cx,cy,cz,len,mass,comx,comy,comz}. */
void orc_geo_export(const orc_geo *g, int32_t *ints, double *dbl, int32_t *child)
{
    for (int64_t k = 0; k < g->nn; ++k) {
        const gnode *q = &g->nodes[k];
        ints[2 * k] = q->level; ints[2 * k + 1] = geo_external(q);
        double *d = dbl + 8 * k;
        d[0] = q->cx; d[1] = q->cy; d[2] = q->cz; d[3] = q->len;
        d[4] = q->com.mass; d[5] = q->com.x; d[6] = q->com.y; d[7] = q->com.z;
        if (child) for (int s = 0; s < 8; ++s) child[8 * k + s] = q->child[s];
    }
}

/* Interaction::singleInteraction(body1, body2, symmetric=false) in SI units
 * (Interaction.cpp:17-55) when kernel == 0:
 *   pd = (p1 - p2) * TO_METERS; distance = |pd|; skip if 0;
 *   force = TIME_STEP * (G * m1 * m2) / ((distance^2 + SOFTENING^2) * distance);
 *   acc1 -= force * pd / m1.
 * kernel == 1 is the Plummer G = 1 form used by the engine (acc1 += m2 d / r^3). */
typedef struct { double tm, g, ts, soft, eps2; int kernel; } geo_units;

static inline void geo_single(const geo_units *u, const gbody *b1, const gbody *b2, double *acc)
{
    if (u->kernel == 0) {
        double pdx = (b1->x - b2->x) * u->tm;
        double pdy = (b1->y - b2->y) * u->tm;
        double pdz = (b1->z - b2->z) * u->tm;
        double distance = sqrt(pdx * pdx + pdy * pdy + pdz * pdz);
        if (distance == 0) return;
        double force = u->ts * (u->g * b1->mass * b2->mass) /
                       ((distance * distance + u->soft * u->soft) * distance);
        acc[0] -= force * pdx / b1->mass;
        acc[1] -= force * pdy / b1->mass;
        acc[2] -= force * pdz / b1->mass;
    } else {
        double dx = b2->x - b1->x, dy = b2->y - b1->y, dz = b2->z - b1->z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 == 0.0) return;
        double r2 = d2 + u->eps2;
        double inv = b2->mass / (r2 * sqrt(r2));
        acc[0] += dx * inv; acc[1] += dy * inv; acc[2] += dz * inv;
    }
}

static void geo_walk(const orc_geo *g, int32_t node, const gbody *b, double theta,
                     const geo_units *u, double *acc, int64_t *cnt)
{
    const gnode *q = &g->nodes[node];
    if (geo_external(q)) {
        geo_single(u, b, &q->com, acc);
        (*cnt)++;
        return;
    }
    double dx = q->com.x - b->x, dy = q->com.y - b->y, dz = q->com.z - b->z;
    if (q->len / sqrt(dx * dx + dy * dy + dz * dz) < theta) {
        geo_single(u, b, &q->com, acc);
        (*cnt)++;
        return;
    }
    for (int s = 0; s < 8; ++s)
        if (q->child[s] != -1) geo_walk(g, q->child[s], b, theta, u, acc, cnt);
}

/* Accelerations of all n bodies from the reference octree (bodies outside the
 * root are skipped and get 0, Interaction.cpp:203-208).  kernel 0: SI reference
 * form with to_meters, G and time_step (returns the reference's velocity
 * increment); kernel 1: Plummer G = 1 with eps2. */
void orc_geo_accel(const orc_geo *g, int64_t n, const double *x, const double *y, const double *z,
                   const double *m, double theta, int kernel, double to_meters, double grav,
                   double time_step, double softening, double eps2, double *acc, int64_t *ninter)
{
    geo_units u = {to_meters, grav, time_step, softening, eps2, kernel};
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        double a[3] = {0.0, 0.0, 0.0};
        int64_t cnt = 0;
        gbody b = {m[i], x[i], y[i], z[i]};
        if (geo_contains(&g->nodes[0], x[i], y[i], z[i])) geo_walk(g, 0, &b, theta, &u, a, &cnt);
        acc[3 * i] = a[0]; acc[3 * i + 1] = a[1]; acc[3 * i + 2] = a[2];
        if (ninter) ninter[i] = cnt;
    }
}

/* ============================================ stars ("suns") outside the tree */
/* The star part of Interaction::interactBodies (Interaction.cpp:111-124) in the
 * order one thread runs it: for each sun s, every other sun ss
 * (singleInteraction(&suns[s], &suns[ss], false): only s is accelerated), then
 * every body b (singleInteraction(&suns[s], &bods[b], true): the one force of
 * the pair accelerates both).  Suns are not in the tree.
 * kernel 0: the reference's SI form (velocity increments, the expressions of
 * Interaction.cpp:17-55); kernel 1: Plummer with G = 1 (accelerations, the
 * engine's form: sun s gets sum m_b (x_b - x_s) r^-3, body b gets
 * sum m_s (x_s - x_b) r^-3, zero distances left out).
 * sun_acc (3 ns) and body_acc (3 nb) are added to. */
void orc_star_forces(int64_t ns, const double *sm, const double *sx, const double *sy, const double *sz,
                     int64_t nb, const double *bm, const double *bx, const double *by, const double *bz,
                     int kernel, double to_meters, double grav, double time_step, double softening,
                     double eps2, double *sun_acc, double *body_acc)
{
    geo_units u = {to_meters, grav, time_step, softening, eps2, kernel};
    for (int64_t s = 0; s < ns; ++s) {
        const gbody a = {sm[s], sx[s], sy[s], sz[s]};
        double *as = sun_acc + 3 * s;
        for (int64_t ss = 0; ss < ns; ++ss) {
            if (ss == s) continue;
            const gbody b = {sm[ss], sx[ss], sy[ss], sz[ss]};
            geo_single(&u, &a, &b, as);
        }
        for (int64_t i = 0; i < nb; ++i) {
            const gbody b = {bm[i], bx[i], by[i], bz[i]};
            double *ab = body_acc + 3 * i;
            if (kernel == 0) {
                const double pdx = (a.x - b.x) * u.tm, pdy = (a.y - b.y) * u.tm, pdz = (a.z - b.z) * u.tm;
                const double distance = sqrt(pdx * pdx + pdy * pdy + pdz * pdz);
                if (distance == 0) continue;
                const double force = u.ts * (u.g * a.mass * b.mass) / ((distance * distance + u.soft * u.soft) * distance);
                as[0] -= force * pdx / a.mass;
                as[1] -= force * pdy / a.mass;
                as[2] -= force * pdz / a.mass;
                ab[0] += force * pdx / b.mass;
                ab[1] += force * pdy / b.mass;
                ab[2] += force * pdz / b.mass;
            } else {
                geo_single(&u, &a, &b, as);
                geo_single(&u, &b, &a, ab);
            }
        }
    }
}

/* Interaction::updateBodies (Interaction.cpp:219-255) for one set of bodies:
 * v += a (the velocity increment of the step), a = 0,
 * x += TIME_STEP * v / TO_METERS. */
void orc_update_reference(int64_t n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                          double *acc, double time_step, double to_meters)
{
    for (int64_t i = 0; i < n; ++i) {
        vx[i] += acc[3 * i]; vy[i] += acc[3 * i + 1]; vz[i] += acc[3 * i + 2];
        acc[3 * i] = 0.0; acc[3 * i + 1] = 0.0; acc[3 * i + 2] = 0.0;
        x[i] += time_step * vx[i] / to_meters;
        y[i] += time_step * vy[i] / to_meters;
        z[i] += time_step * vz[i] / to_meters;
    }
}
