/*
 * nbx.h -- C ABI of the nbx Barnes-Hut engine for AMD Instinct MI355X (gfx950).
 *
 * The seven stage entry points replace the methods of the reference's launch
 * wrapper class KernelsWrapper (MichaelSt98/NNS 3D/CUDA/CUDA_NBody,
 * include/KernelsWrapper.cuh:28-48, src/KernelsWrapper.cu:80-208).  They take
 * the same device pointers, in the same order, over the same array layout
 * (src/BarnesHut.cu:62-102):
 *
 *   float  x, y, z, mass, vx, vy, vz, ax, ay, az   [m]   bodies in [0,n), cells in [n,*index)
 *   int    child                                   [8m]  child[8*node+slot]; -1 empty;
 *                                                        the root's slots are child[0..7]
 *   int    count, start, sorted                    [m]
 *   int    index, mutex                            scalars in device memory
 *   float  minX, maxX, minY, maxY, minZ, maxZ      scalars in device memory
 *
 * The caller owns every buffer above (the reference allocates them in
 * BarnesHut::BarnesHut, src/BarnesHut.cu:62-102).  What the engine needs
 * beyond the reference layout (keys, sort buffers, FP64 cell sums, the
 * traversal layout) lives in an nbx_context that is allocated once by
 * nbx_create and never inside a stage.
 *
 * Differences from the reference, all deliberate:
 *   - every stage takes a stream (hipStream_t passed as void*, NULL = the null
 *     stream) and an optional elapsed_ms out-parameter; when elapsed_ms is not
 *     NULL the stage is bracketed by hipEvents on that stream and waits for the
 *     second event, like the reference's `timing` flag (KernelsWrapper.cu:84-99);
 *   - stages return an int status (NBX_OK or a negative NBX_ERR_*) instead of
 *     nothing; errors raised inside kernels (tree capacity, bucket overflow)
 *     are latched in the context and reported by nbx_check; once latched, the
 *     device stages that follow on that context, nbx_update included (the
 *     bodies stay where they are), do no work until the next nbx_reset_arrays,
 *     which keeps the latched error for nbx_check to report;
 *   - theta, softening, G, dt and d are run-time parameters (nbx_params); the
 *     reference hard-codes theta = 1.5 and eps^2 = 0.0025 (src/Kernels.cu:17-18)
 *     and ignores its `g` argument;
 *   - cells are Morton-key prefixes: child slot = x + 2y + 4z of the key digit,
 *     1 for the upper half (3D/OpenMP/OpenMP_NBody/src/Body.cpp:29); the
 *     reference's CUDA childPath is the complement (src/Kernels.cu:198-222);
 *   - the opening test is the reference CPU code's per-particle l/D < theta
 *     (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:64-67) with a Plummer
 *     kernel a = G m d (|d|^2 + eps^2)^-3/2.
 * The stages must be called in the reference's order (reset, bounding box,
 * build, centre of mass, sort, forces, update) on one context; nbx_build_tree
 * binds the tree arrays it is given to the context for the stages after it.
 * A context's calls must not overlap on the device: its stages share the
 * context's buffers (keys, tree scratch, the force launch's job record and
 * work counters), so a context is used from one stream at a time (work queued
 * on another stream must be ordered after it by the caller); independent
 * systems that run concurrently take a context each.
 */
#ifndef NBX_H
#define NBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_OK 0
#define NBX_ERR_INVALID (-1)   /* bad argument (null pointer, n > n_max, m too small, ...) */
#define NBX_ERR_CAPACITY (-2)  /* more cells needed than m - n                            */
#define NBX_ERR_BUCKET (-3)    /* more than 8 bodies share one 63-bit key                */
#define NBX_ERR_STACK (-4)     /* traversal stack overflow (cannot happen for depth <= 21) */
#define NBX_ERR_NONFINITE (-5) /* NaN/Inf position or a zero/negative total mass          */
#define NBX_ERR_HIP (-6)       /* HIP runtime error, or the summary's workgroups could not
                                  all run at once (its bounded grid-barrier wait ran out);
                                  see nbx_last_error                                     */
#define NBX_ERR_ORDER (-7)     /* stage called before the stage it depends on            */

#define NBX_KEY_MORTON 0
#define NBX_KEY_HILBERT 1

typedef struct nbx_context nbx_context;

/* Run-time parameters (the reference's SimulationParameters, include/Constants.h:4-22,
 * plus the constants it hard-codes in src/Kernels.cu:17-18). */
typedef struct nbx_params {
    float theta;      /* opening angle: a cell is accepted when l/D < theta            */
    float softening;  /* Plummer epsilon; eps^2 is added to |d|^2                       */
    float gravity;    /* G                                                              */
    float timestep;   /* dt                                                             */
    float dampening;  /* d in x += d dt v                                               */
    int kahan;        /* 1: compensated (Kahan) accumulation of the accelerations       */
    int block_size;   /* threads per block of the update stage: 0 (= 256) or a multiple  */
                      /* of 64 in [64, 1024] (the reference's blockSize, main.cu:30)   */
    int force_waves;  /* waves per block of the force kernel: 2, 4 or 8; 0 = 4 in
                         nbx_create, "keep the current choice" in nbx_set_params        */
    int build_locked; /* 1: subtrees built by lock-based insertion (compare-and-swap on
                         child slots, Kernels.cu:138-241); 0: from the shared-digit
                         counts of neighbouring sorted keys, one store per slot.  Both
                         give the same cells and child slots.                          */
} nbx_params;

/* Fills p with the engine defaults: theta 0.5, softening 1e-3, G 1, dt 1e-3, d 1, Kahan on. */
void nbx_default_params(nbx_params *p);

/* Creates a context for up to n_max bodies and m_max nodes (the reference uses
 * m = 2n + 12000, src/BarnesHut.cu:18).  device: HIP device ordinal the context
 * allocates on and launches on.  n_max is at most 48,805,888 (NBX_ERR_INVALID
 * above): the tree's top table holds 2^20 entries, 22 per 1024 bodies.  m_max is
 * at most 2^27 = 134,217,728: the force walk's node descriptors hold a frame
 * record offset (or a sorted position) times 16 in 32 bits. */
int nbx_create(const nbx_params *p, int n_max, int m_max, int device, nbx_context **out);
int nbx_destroy(nbx_context *ctx);
int nbx_set_params(nbx_context *ctx, const nbx_params *p);
int nbx_get_params(const nbx_context *ctx, nbx_params *p);
/* Message of the last error on this context ("" if none).  ctx may be NULL for nbx_create errors. */
const char *nbx_last_error(const nbx_context *ctx);
/* Waits for the stream, then returns NBX_OK or a device-side error latched since the
 * last nbx_check: an error of an earlier step is kept by nbx_reset_arrays and
 * reported once here, so a caller may check once after many steps. */
int nbx_check(nbx_context *ctx, void *stream);

/* ---- the seven stages (KernelsWrapper.cuh:28-48) ------------------------------ */

/* resetArrays (KernelsWrapper.cu:80-106, Kernels.cu:21-64): child = -1; count = 1 for
 * bodies, 0 for cells; cell x,y,z,mass = 0; start = -1; sorted = 0; *index = n;
 * *mutex = 0; bounding box = 0; clears the context's error latch (an error latched
 * there stays reported by nbx_check). */
int nbx_reset_arrays(nbx_context *ctx, int *mutex, float *x, float *y, float *z, float *mass,
                     int *count, int *start, int *sorted, int *child, int *index, float *minX,
                     float *maxX, float *minY, float *maxY, float *minZ, float *maxZ, int n, int m,
                     void *stream, float *elapsed_ms);

/* computeBoundingBox (KernelsWrapper.cu:108-134, Kernels.cu:67-154): min/max of the n
 * bodies (the reference also folds in the origin through its reset to 0; this stage does
 * not), plus the FP64 extent (xMin, yMin, zMin, maxSpan) that fixes key quantisation. */
int nbx_compute_bounding_box(nbx_context *ctx, int *mutex, float *x, float *y, float *z,
                             float *minX, float *maxX, float *minY, float *maxY, float *minZ,
                             float *maxZ, int n, void *stream, float *elapsed_ms);

/* buildTree (KernelsWrapper.cu:136-162, Kernels.cu:157-377): x, y, z and n must be those of
 * the last nbx_compute_bounding_box (NBX_ERR_INVALID otherwise).  Morton keys, stable radix
 * sort of (key, body), then lock-based insertion along the key digits (child slots
 * locked with -2 by compare-and-swap).  Fills child and *index.  Cell sums are left to
 * nbx_centre_of_mass. */
int nbx_build_tree(nbx_context *ctx, float *x, float *y, float *z, float *mass, int *count,
                   int *start, int *child, int *index, float *minX, float *maxX, float *minY,
                   float *maxY, float *minZ, float *maxZ, int n, int m, void *stream,
                   float *elapsed_ms);

/* centreOfMass (KernelsWrapper.cu:164-176, Kernels.cu:381-399): deterministic bottom-up
 * summary.  Cell mass and mass moments are summed in FP64 over the children in slot
 * order; writes cell x,y,z (centre of mass), mass, count and start. */
int nbx_centre_of_mass(nbx_context *ctx, float *x, float *y, float *z, float *mass, int *index,
                       int n, void *stream, float *elapsed_ms);

/* sort (KernelsWrapper.cu:178-190, Kernels.cu:403-456): sorted[k] = k-th body in key
 * order (equal to the reference's in-order tree placement for this slot order), and
 * the traversal layout used by nbx_compute_forces. */
int nbx_sort(nbx_context *ctx, int *count, int *start, int *sorted, int *child, int *index, int n,
             void *stream, float *elapsed_ms);

/* computeForces (KernelsWrapper.cu:192-206, Kernels.cu:460-577): a_i = G sum m d (|d|^2 +
 * eps^2)^-3/2 over body i's per-particle interaction list; written to ax,ay,az[i]. */
int nbx_compute_forces(nbx_context *ctx, float *x, float *y, float *z, float *vx, float *vy,
                       float *vz, float *ax, float *ay, float *az, float *mass, int *sorted,
                       int *child, float *minX, float *maxX, int n, float g, void *stream,
                       float *elapsed_ms);

/* update (KernelsWrapper.cu:208, Kernels.cu:581-600): v = fma(dt, a, v); x = fma(d*dt, v, x).
 * Does nothing while an error of this step is latched (see nbx_check). */
int nbx_update(nbx_context *ctx, float *x, float *y, float *z, float *vx, float *vy, float *vz,
               float *ax, float *ay, float *az, int n, float dt, float d, void *stream,
               float *elapsed_ms);

/* ---- extensions (no counterpart in KernelsWrapper) ------------------------------ */

/* Keys of n bodies for the extent found by the last nbx_compute_bounding_box (over the
 * same x, y, z and n, else NBX_ERR_INVALID):
 * NBX_KEY_MORTON is Body::getKey (3D/OpenMP/OpenMP_NBody/src/Body.cpp:13-32);
 * NBX_KEY_HILBERT is Lebesgue2Hilbert of its 63 key bits
 * (MolecularDynamics/BarnesHutParallel/SpaceFillingCurves/src/Tree.cpp:82-104). */
int nbx_compute_keys(nbx_context *ctx, const float *x, const float *y, const float *z, int n,
                     int key_type, uint64_t *keys, void *stream, float *elapsed_ms);

/* Stable ascending LSD radix sort of n 64-bit keys: sorted_keys and perm (perm[k] = index of
 * the k-th key).  Scratch comes from the context (n <= n_max). */
int nbx_sort_keys(nbx_context *ctx, const uint64_t *keys, int n, uint64_t *sorted_keys,
                  int32_t *perm, void *stream, float *elapsed_ms);

/* Stars ("suns") outside the tree: the reference's OpenMP step keeps its suns apart from
 * the bodies it inserts into the octree and couples them directly, both ways
 * (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:111-124; updated with the bodies,
 * :219-234).  This stage adds to each of the n bodies' ax, ay, az (call it after
 * nbx_compute_forces, which writes them) the pull of the ns stars, and writes the
 * stars' own accelerations sax, say, saz: the pull of every other star and of every
 * body.  Plummer kernel G m d (|d|^2 + eps^2)^-3/2 with the context's softening; zero
 * distances are left out.  The terms are summed in FP64 in a fixed order (the same
 * result on every run).  0 <= ns <= NBX_MAX_STARS; the stars' arrays are device
 * arrays of ns floats.  Does nothing while an error of this step is latched. */
#define NBX_MAX_STARS 8
int nbx_star_forces(nbx_context *ctx, const float *sx, const float *sy, const float *sz, const float *smass,
                    float *sax, float *say, float *saz, int ns, const float *x, const float *y, const float *z,
                    const float *mass, float *ax, float *ay, float *az, int n, float g, void *stream,
                    float *elapsed_ms);

/* The sort of nbx_sort_keys (stable ascending LSD radix sort of n 64-bit keys into
 * sorted_keys and perm) without a context and without allocating: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory -- code under HIP
 * graph capture, serving loops.  Its scratch is one static device array of the library
 * (12.5 MiB on every GPU it runs on, held until the process ends), found on the current
 * device at the first call there; n is at most 1,048,576 (NBX_ERR_INVALID above, message
 * in nbx_last_error(NULL)).  keys, sorted_keys and perm are device arrays on the current
 * device; stream (NULL = the null stream) must belong to it.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int nbx_sort_keys_static(const uint64_t *keys, int n, uint64_t *sorted_keys, int32_t *perm, void *stream);

/* Per-body interaction counts of the last tree (cells accepted plus leaf bodies other
 * than the body itself), written to counts[i] for body i.  Not part of a time step. */
int nbx_count_interactions(nbx_context *ctx, int n, int64_t *counts, void *stream);

/* Work of the force walk over the last tree (not part of a time step): out[0] the
 * interactions (the sum of the per-particle lists, as nbx_count_interactions),
 * out[1] the child steps (one step tests one child for the 64 lanes of a group, so
 * 64 out[1] lane slots hold out[0] useful ones), out[2] the frames (opened cells
 * walked), out[3] the groups of 64 bodies.  The reference's warp-vote walk pays the
 * same way for the union of a warp's lists (Kernels.cu:543). */
int nbx_walk_stats(nbx_context *ctx, int64_t *out, void *stream);

/* One full step (reset, bounding box, build, centre of mass, sort, forces, update) over
 * the arrays of the reference layout; stage_ms (7 floats, nullable) receives per-stage
 * hipEvent times without host waits between stages.  With nstars > 0 the step is the
 * reference's OpenMP step with suns (Interaction::interactBodies): nbx_star_forces runs
 * after the tree forces (its time counts in the force stage) and the stars are updated
 * with the bodies (in the update stage). */
typedef struct nbx_arrays {
    float *x, *y, *z, *mass, *vx, *vy, *vz, *ax, *ay, *az;
    int *count, *start, *sorted, *child, *index, *mutex;
    float *minX, *maxX, *minY, *maxY, *minZ, *maxZ;
    int n, m;
    /* stars outside the tree: nstars (0..NBX_MAX_STARS) device arrays of each */
    float *sx, *sy, *sz, *smass, *svx, *svy, *svz, *sax, *say, *saz;
    int nstars;
} nbx_arrays;
int nbx_step(nbx_context *ctx, const nbx_arrays *a, void *stream, float *stage_ms);

/* Copies a context buffer of the last tree to host memory (diagnostics and tests).
 * count is in elements of the buffer's type. */
#define NBX_BUF_KEYS 1        /* uint64 [n] key of body i                        */
#define NBX_BUF_SORTED_KEYS 2 /* uint64 [n] keys in sorted order                 */
#define NBX_BUF_PERM 3        /* int32  [n] body of sorted position k            */
#define NBX_BUF_CELL_LEVEL 4  /* int32  [m] level of cell c (root 0), -1 unused  */
#define NBX_BUF_CELL_MASS64 5 /* double [m] FP64 mass of cell c                  */
#define NBX_BUF_CELL_COM64 6  /* double [3m] FP64 centre of mass of cell c       */
#define NBX_BUF_ROOT 7        /* double [4] root mass, com x, y, z               */
#define NBX_BUF_EXTENT 8      /* double [4] xMin, yMin, zMin, maxSpan            */
int nbx_copy_buffer(nbx_context *ctx, int which, void *host_dst, int64_t count, void *stream);

/* Times the force kernel at 2, 4 and 8 waves per block on the current tree (after nbx_sort)
 * and keeps the fastest for this context's later nbx_compute_forces calls.  *waves_out
 * (nullable) receives the choice.  Synchronises the stream. */
int nbx_autotune(nbx_context *ctx, void *stream, int *waves_out);

/* ---- differentiation (no counterpart in the reference) ----------------------------- */

/* Vector-Jacobian product of the accelerations over the tree of the last nbx_sort: for
 * u = dL/da (device arrays of n floats, input order), writes dL/dx, dL/dy, dL/dz and
 * dL/dm of every body (n floats each).  g is the gravity given to nbx_compute_forces.
 * Each body meets exactly the nodes of its forward interaction list; cell gradients
 * flow down to the bodies through the centre-of-mass sums (FP64).  The node terms
 * are logged, sorted by node and summed; the order of records within a node
 * follows the log, so the last bits may differ between runs.  The first call
 * allocates the scratch: 64 (n_max + m_max + 8) bytes plus 40 bytes per log record, with
 * 64 n_max + 4 n_max records (about 2.6 KB per body); if a call needs more
 * records, the rest are added with atomics. */
int nbx_force_vjp(nbx_context *ctx, const float *ux, const float *uy, const float *uz, float g, float *gx,
                  float *gy, float *gz, float *gm, void *stream, float *elapsed_ms);

/* ---- batched ensembles (no counterpart in the reference) --------------------------- */

/* nsys independent systems in one set of concatenated arrays: system s owns bodies
 * [off_s, off_s + sys_n[s]) with off_s = sys_n[0] + ... + sys_n[s-1].  sys_n and sys_theta
 * (per-system opening angle; NULL = p->theta) are host arrays.  Every stage is one launch
 * for the whole ensemble; the tree of system s hangs off its own root cell.  At most
 * 65535 systems and 47,662 samples: the sum over systems of ceil(sys_n[s] / 1024). */
typedef struct nbx_batch nbx_batch;
int nbx_batch_create(const nbx_params *p, int nsys, const int *sys_n, const float *sys_theta, int device,
                     nbx_batch **out);
int nbx_batch_destroy(nbx_batch *b);
const char *nbx_batch_last_error(const nbx_batch *b);
/* Binds the caller's concatenated device arrays (N = sum of sys_n elements each). */
int nbx_batch_bind(nbx_batch *b, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax,
                   float *ay, float *az, float *mass);
/* One step of every system (reset, bounding boxes, build, summary, layout, forces, update);
 * stage_ms (7 floats, nullable) as in nbx_step. */
int nbx_batch_step(nbx_batch *b, void *stream, float *stage_ms);
/* Same step, captured once into a hipGraph on `stream` (not NULL) and replayed. */
int nbx_batch_graph_step(nbx_batch *b, void *stream);
int nbx_batch_check(nbx_batch *b, void *stream);
/* Per-body interaction counts of the last tree, device array of N int64. */
int nbx_batch_count_interactions(nbx_batch *b, int64_t *counts, void *stream);
/* nbx_walk_stats of the ensemble's last tree (groups: the force table's groups). */
int nbx_batch_walk_stats(nbx_batch *b, int64_t *out, void *stream);
/* nbx_autotune for the ensemble's force kernel (invalidates a captured graph). */
int nbx_batch_autotune(nbx_batch *b, void *stream, int *waves_out);
/* Host copy of a batch buffer: NBX_BUF_KEYS, NBX_BUF_SORTED_KEYS, NBX_BUF_PERM (N each),
 * NBX_BUF_EXTENT (4 per system), NBX_BUF_CELL_MASS64 / NBX_BUF_CELL_COM64 / NBX_BUF_CELL_LEVEL
 * (node index space, roots at N + s). */
int nbx_batch_copy_buffer(nbx_batch *b, int which, void *host_dst, int64_t count, void *stream);

/* Seeded initial conditions written to host arrays of length n (positions, velocities,
 * masses; the rest untouched).  The generator is xoshiro256** seeded by splitmix64. */
int nbx_ic_star_disk(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy,
                     float *vz, float *mass);
int nbx_ic_plummer(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy,
                   float *vz, float *mass);
int nbx_ic_binary_disk(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy,
                       float *vz, float *mass);

/* ---- diagnostics (no counterpart in the reference) --------------------------------- */

/* Issue-rate probe of the FP32 vector ALU on the current device: `blocks` blocks of 256
 * threads, each lane `iters` rounds of 8 independent instructions, kind 0 = v_fma_f32
 * (2 flop each), kind 1 = v_rsq_f32; 1 <= blocks <= 2^22.  out: device array of out_count >= 256 * blocks
 * floats (one result per lane).  *ms receives the kernel time (hipEvents on `stream`;
 * synchronises it). */
int nbx_probe_valu(int kind, int blocks, int iters, float *out, int64_t out_count, void *stream, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* NBX_H */
