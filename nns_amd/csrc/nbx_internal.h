// nbx_internal.h -- context layout and kernel launchers shared by the nbx sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/nbx.h"

namespace nbx {

// Key geometry: 21 levels of 3 bits (Body.h:32 MASK_COORD2KEY = 2^21 - 1).
constexpr int kKeyLevels = 21;
constexpr double kMaskCoord = 2097151.0;

// Radix sort: 8-bit digits, 256 threads x 8 items per tile.  (9-bit digits would
// save a pass on a 63-bit Morton key, but their 512-way scatter is slower: 10
// passes of 9 bits took longer than 12 of 8 at N = 4M.)
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;

// Force traversal: one body per lane, 64-body groups, depth-first stack per wave.
// A cell at level L is opened at most once per path and pushes at most 8 children,
// so the stack never holds more than 8 + 7*21 = 155 entries.
constexpr int kForceStack = 160;
constexpr int kRootRef = 0x7fffffff;

// Device-side error latch bits (mapped to NBX_ERR_* on the host).
constexpr int kErrCapacity = 1;
constexpr int kErrBucket = 2;
constexpr int kErrStack = 4;
constexpr int kErrNonFinite = 8;
constexpr int kErrStall = 16;  // a grid barrier's bounded wait ran out (workgroups not resident together)

// Tree construction: cells holding more than kSubtreeMax bodies form the "top"
// of the tree and are derived from the sorted keys; every child range of a top
// cell with 2..kSubtreeMax bodies is a subtree built by one workgroup.
constexpr int kSubtreeMax = 1024;
constexpr int kBboxChunk = 4096;  // bodies per workgroup of the ensemble's bounding box
constexpr int kLevels = kKeyLevels + 1;  // levels 0..21 per sample in the top table
// The top table (kLevels entries per sample of kSubtreeMax bodies) is scanned by
// a two-level scan of at most 2^20 entries: the largest system, or ensemble
// (each system's samples rounded up), is 47,662 samples = 48,805,888 bodies.
constexpr int kMaxSamples = (1 << 20) / kLevels;
constexpr int kMaxBodies = kMaxSamples * kSubtreeMax;
// A frame descriptor's tag holds a record offset (or a sorted position) times 16
// in an int: nodes, and so records, stay below 2^27.
constexpr int kMaxNodes = 1 << 27;
// Frame blocks are packed: block b's children are records [off(b), off(b) + k_b)
// of fnode / fmeta.  off(b) is summed over runs of kOffRun blocks, then over the
// runs (frame_count_kernel, then a scan of the run totals).
constexpr int kOffRun = 1024;

// Arguments of a force launch that a group needs only when its walk starts or
// ends.  A setup kernel stores them in device memory before the launch and the
// walk reads them there, so that they hold no scalar registers during the walk.
struct ForceJob {
    const float4 *nodes4;   // bodies (x, y, z, m) by sorted position
    const int32_t *gorder;  // group order: sorted positions
    const float4 *froot4;   // root of each system as a child (x, y, z, m)
    const int2 *frootm;     // and its descriptor
    const int4 *groups;     // ensemble: {first, count, system} per group
    const int32_t *perm;    // sorted position -> body
    float *ax, *ay, *az;    // accelerations by body
    int64_t *counts;        // counting variant: interactions per body
    int64_t *walk;          // counting variant: child steps and frames per group (nullable)
    int32_t *gcost;         // child steps per group, written by every launch
    int *err;               // error latch
    int n, ngroups;
    float G;
};

struct DeviceState {
    double ext[4];      // xMin, yMin, zMin, maxSpan
    double root64[4];   // root mass, com x, y, z
    float4 root4;       // root com x, y, z, mass (float)
    int err;            // latched kErr* bits of the current step (cleared by the reset stage)
    int sticky;         // unreported bits of earlier steps, moved here by the reset stage
    int seen;           // bits of err already reported by the host check (not moved to sticky)
    int nsub;           // subtrees found by the top-tree pass
    int ntop;           // top cells (excluding roots): ids [base0, base0 + ntop)
    int log_next;       // adjoint: next free record of the node-term log
    int log_count;      // adjoint: records written (sort length)
    // Grid barrier of the top summary: arrivals of its workgroups and their
    // exits; the last workgroup to leave sets both back to 0.
    alignas(128) int tsync;
    int tdone;
    alignas(128) ForceJob fj;  // the current force launch's job record
    int fcut[9];               // its groups per XCD: XCD x walks [fcut[x], fcut[x + 1])
    int fcost_n;               // groups of the launch that wrote ForceJob::gcost last
};

// Tree-construction scratch shared by the single-system and ensemble paths.
struct TreeScratch {
    uint8_t *lcp = nullptr;      // [n-1] common key digits of sorted neighbours
    uint32_t *top_flag = nullptr, *top_id = nullptr;  // [samples * kLevels]
    int32_t *top_a = nullptr, *top_b = nullptr;       // [samples * kLevels] body range of a top cell
    int4 *sub = nullptr;         // {first, end, level, slot address}
    int4 *subcells = nullptr;    // {first cell id, cells, level, deepest level} of each subtree
    uint32_t *cbase = nullptr;   // [sub_max + 1] cells of each subtree, then their exclusive sum
    uint32_t *scan_tmp = nullptr;  // 1024 block sums of the top-table scan
    int samples_max = 0, sub_max = 0;
};

// Host copy of the FP64 cell sums as NBX_BUF_CELL_MASS64 (count masses) or, com,
// NBX_BUF_CELL_COM64 (count coordinates, three per cell: the moments divided by
// the mass, 0 for a massless cell).
hipError_t copy_cell_sums(const double4 *sum64, bool com, double *dst, size_t count);

// Message returned by nbx_last_error(NULL) (errors of calls without a context).
void set_create_error(const std::string &msg);

// Allocates the scratch for up to n bodies in nseg segments with `samples` top-table
// samples in total; returns false on allocation failure.
bool alloc_tree_scratch(TreeScratch &ts, int n, int nseg, int samples);
void free_tree_scratch(TreeScratch &ts);

}  // namespace nbx

struct nbx_context {
    nbx_params p{};
    int device = 0;
    int n_max = 0, m_max = 0;
    std::string last_error;

    // context-owned device memory
    nbx::DeviceState *st = nullptr;
    float *bbox_part = nullptr;  // 6 floats per bbox block
    int bbox_blocks = 0;
    uint64_t *keys = nullptr, *skeys = nullptr, *kbuf = nullptr;
    int32_t *perm = nullptr, *vbuf = nullptr, *pos = nullptr;
    uint32_t *hist = nullptr;
    int hist_tiles = 0;
    int32_t *cell_level = nullptr;
    double4 *sum64 = nullptr;    // [m_max] FP64 sums of each cell: mass, mass moments x, y, z
    float4 *nodes4 = nullptr;
    float4 *fnode = nullptr;     // [m_max + 8] frame records of the force walk: child (x, y, z, m)
    int2 *fmeta = nullptr;       // [m_max + 8] child descriptors
    uint32_t *boff = nullptr;    // [m_max + 1] record offset of each block within its run of kOffRun
    uint32_t *bsum = nullptr;    // [m_max / kOffRun + 1] records before each run
    float4 *froot4 = nullptr;    // root as a child of a virtual parent
    int2 *frootm = nullptr;
    int32_t *pblk = nullptr;     // [m_max] parent block of each block (-1: root)
    int32_t *bpar = nullptr;     // [n_max] parent block of each body, by sorted position
    // adjoint scratch (allocated by the first nbx_force_vjp)
    double *adj_tgt = nullptr;   // [4 (n_max + m_max + 1)] gradient sums of bodies (sorted), then records
    double *adj_ab = nullptr;    // [4 (m_max + 1)] per block: sum over ancestors of G/M and GM - G.X/M
    // node-term log: one record {node, float4} per child step with accepting lanes
    int64_t adj_cap = 0;         // records
    uint32_t *adj_log_id = nullptr, *adj_sk = nullptr, *adj_tk = nullptr;
    float4 *adj_log_val = nullptr;
    int32_t *adj_sv = nullptr, *adj_tv = nullptr;
    uint32_t *adj_hist = nullptr;
    float layout_theta = -1.f;   // theta the frame blocks' thresholds were computed with
    float layout_eps2 = -1.f;    // and eps^2
    int64_t *counts = nullptr;
    float *acc_scratch = nullptr;
    double *star_part = nullptr;  // star forces: per-block partial pulls, kStarBlocks x NBX_MAX_STARS x 3
    float4 *bodies4 = nullptr;  // (x, y, z, m) of every body in input order
    uint32_t *gkey = nullptr, *gkey2 = nullptr;  // group-order sort keys
    int32_t *gorder = nullptr;  // [n] sorted position of the j-th body of the force groups
    nbx::TreeScratch ts;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t stage_ev[8] = {};

    // tree bound by nbx_build_tree for the later stages
    float *x = nullptr, *y = nullptr, *z = nullptr, *mass = nullptr;
    int *child = nullptr, *count = nullptr, *start = nullptr, *index = nullptr;
    int n = 0, m = 0;
    int stage = 0;  // 0 reset, 1 bbox, 2 built, 3 summarised, 4 sorted, 5 forces
    // bodies the FP64 extent of the last nbx_compute_bounding_box was taken over
    const float *bbox_x = nullptr, *bbox_y = nullptr, *bbox_z = nullptr;
    int bbox_n = -1;
    bool fresh = false;  // arrays reset since the last build (root slots are -1)
    int force_waves = 4; // waves per block of the force kernel (2, 4 or 8; autotuned)
    int64_t *walk = nullptr;     // [2 ceil(n_max / 64)] child steps and frames per force group
    int32_t *gcost = nullptr;    // [ceil(n_max / 64)] child steps per force group of the last launch
};

// Ensemble of S independent systems (batched form).  System s owns bodies
// [off[s], off[s+1]) of the concatenated arrays and the root cell N + s; the
// other cells are allocated from N + S upwards.  All arrays are owned by the
// batch except the bound body arrays.
struct nbx_batch {
    nbx_params p{};
    int device = 0;
    int S = 0, N = 0, m = 0;
    int ngroups = 0;
    std::string last_error;
    int force_waves = 4;
    int64_t *walk = nullptr;  // [2 ngroups]
    int32_t *gcost = nullptr; // [ngroups]
    // device
    nbx::DeviceState *st = nullptr;
    int32_t *off = nullptr, *sys_of = nullptr;
    float *theta = nullptr;
    int4 *groups = nullptr;  // {first sorted position, bodies, system, 0}
    double *ext = nullptr;   // 4 per system
    float *mm = nullptr;     // 6 per system
    int nchunk = 0;                 // bounding-box chunks: at most kBboxChunk bodies of one system each
    int4 *bchunk = nullptr;         // [nchunk] {system, first body, end, 0}
    int32_t *bchunk_off = nullptr;  // [S+1] first chunk of each system
    float *bpart = nullptr;         // [6 nchunk] partial extremes
    uint64_t *keys = nullptr, *skeys = nullptr, *kbuf = nullptr, *k2 = nullptr, *k2buf = nullptr;
    int32_t *perm = nullptr, *perm1 = nullptr, *vbuf = nullptr, *pos = nullptr, *sorted = nullptr;
    uint32_t *hist = nullptr;
    int *child = nullptr, *index = nullptr, *count = nullptr, *start = nullptr;
    int32_t *cell_level = nullptr;
    double4 *sum64 = nullptr;    // [M] FP64 sums of each cell: mass, mass moments x, y, z
    float4 *nodes4 = nullptr;
    float4 *fnode = nullptr;     // [M + 8] frame records
    int2 *fmeta = nullptr;
    uint32_t *boff = nullptr;    // [M] record offset of each block within its run of kOffRun
    uint32_t *bsum = nullptr;    // [M / kOffRun + 1] records before each run
    float4 *froot4 = nullptr;    // [S] roots
    int2 *frootm = nullptr;
    int32_t *pblk = nullptr, *bpar = nullptr;
    int64_t *counts = nullptr;
    float *acc_scratch = nullptr;
    float4 *bodies4 = nullptr;
    uint64_t *gkey = nullptr, *gkey2 = nullptr;
    int32_t *gorder = nullptr;
    int32_t *samp_off = nullptr;  // [S+1] first top-table sample of each system
    nbx::TreeScratch ts;
    hipEvent_t stage_ev[8] = {};
    // bound body arrays (caller-owned, N each)
    float *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    float *ax = nullptr, *ay = nullptr, *az = nullptr, *mass = nullptr;
    // captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace nbx {

// launchers (k_basic.hip)
void launch_reset(nbx_context *c, hipStream_t s, int *mutex, float *x, float *y, float *z, float *mass,
                  int *count, int *start, int *sorted, int *child, int *index, float *minX, float *maxX,
                  float *minY, float *maxY, float *minZ, float *maxZ, int n, int m);
void launch_bbox(nbx_context *c, hipStream_t s, const float *x, const float *y, const float *z,
                 float *minX, float *maxX, float *minY, float *maxY, float *minZ, float *maxZ, int n);
// Also packs (x, y, z, mass) into c->bodies4 when mass is given.
void launch_keys(nbx_context *c, hipStream_t s, const float *x, const float *y, const float *z, const float *mass,
                 int n, int key_type, uint64_t *keys);
// Does nothing once the error latch of st is raised (st may be null: always runs).
void launch_update(hipStream_t s, int block, const DeviceState *st, float *x, float *y, float *z, float *vx, float *vy,
                   float *vz, const float *ax, const float *ay, const float *az, int n, float dt, float d);
void launch_fill_f32(hipStream_t s, float *p, float v, int64_t n);
// Stars outside the tree (nbx_star_forces); scratch c->star_part.
constexpr int kStarPartBlocks = 1024;
void launch_star_forces(nbx_context *c, hipStream_t s, const float *sx, const float *sy, const float *sz,
                        const float *sm, float *sax, float *say, float *saz, int ns, const float *x, const float *y,
                        const float *z, const float *mass, float *ax, float *ay, float *az, int n, float g);

// launchers (k_radix.hip)
// Sorts keys_in (not modified) by their low key_bits bits into keys_out / perm_out (stable);
// writes the inverse permutation to inv_out when not null.  tmp_k / tmp_v: scratch of n
// elements.  Morton keys need 63 bits (bit 63 is set in every key).
void launch_radix_sort(nbx_context *c, hipStream_t s, const uint64_t *keys_in, int n, uint64_t *keys_out,
                       int32_t *perm_out, uint64_t *tmp_k, int32_t *tmp_v, int32_t *inv_out, int key_bits);
constexpr int kMortonSortBits = 63;
// exclusive scan of total <= 2^20 values; sums: scratch of 1024 values
void launch_exclusive_scan_u32(uint32_t *data, int total, uint32_t *sums, hipStream_t s);
// The same in one launch of one workgroup (short arrays).
void launch_scan_one_block(uint32_t *data, int total, hipStream_t s);
// npasses passes of kRadixBits bits from bit first_shift; vals_in == nullptr means 0, 1, 2, ...
// n_dev (nullable): device-side count; only the first min(n, *n_dev) keys are sorted
// (grids are sized for n).  tmp_k may alias keys_in when npasses is odd.  K: uint64_t
// or uint32_t.
template <class K>
void launch_radix_passes(uint32_t *hist, hipStream_t s, const K *keys_in, const int32_t *vals_in, int n,
                         int first_shift, int npasses, K *keys_out, int32_t *vals_out, K *tmp_k,
                         int32_t *tmp_v, int32_t *inv_out, const int *n_dev = nullptr);

// launchers (k_tree.hip)
void launch_build(nbx_context *c, hipStream_t s, int n, int m, int *child, int *index);
void launch_summarize(nbx_context *c, hipStream_t s);
void launch_layout(nbx_context *c, hipStream_t s, int *sorted);
// Frame blocks of the force walk for the current theta and softening (part of
// launch_layout); their thresholds hold crit_L + eps^2.
void launch_frame_layout(nbx_context *c, hipStream_t s);
// eps^2 as the walks add it (the same FP32 product everywhere).
inline float layout_eps2(const nbx_params &p) { return p.softening * p.softening; }
inline bool frame_layout_stale(const nbx_context *c)
{
    return c->layout_theta != c->p.theta || c->layout_eps2 != layout_eps2(c->p);
}
// Hilbert digits kept for the force groups' order (top levels of the key).
constexpr int kGroupLevels = 10;

// launchers (k_force.hip)
// counts (nullable): per-body interaction counts; walk (nullable, with counts): per
// group, the child steps and frames of its walk.
void launch_force(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int64_t *counts,
                  int64_t *walk = nullptr);
void launch_force_waves(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int waves);

// batched form (k_basic.hip, k_radix.hip, k_tree.hip, k_force.hip)
void launch_batch_reset(nbx_batch *b, hipStream_t s);
void launch_batch_bbox(nbx_batch *b, hipStream_t s);
void launch_batch_keys(nbx_batch *b, hipStream_t s);
void launch_batch_sort(nbx_batch *b, hipStream_t s);
void launch_batch_build(nbx_batch *b, hipStream_t s);
void launch_batch_summarize(nbx_batch *b, hipStream_t s);
void launch_batch_layout(nbx_batch *b, hipStream_t s);
void launch_batch_force(nbx_batch *b, hipStream_t s, int64_t *counts, int waves, int64_t *walk = nullptr);

// launchers (k_adjoint.hip)
void launch_vjp(nbx_context *c, hipStream_t s, const float *ux, const float *uy, const float *uz, float *gx,
                float *gy, float *gz, float *gm, float g);
// Adjoint scratch of a context (log capacity from n_max; sort passes from m_max).
bool alloc_vjp_scratch(nbx_context *c);
void free_vjp_scratch(nbx_context *c);

}  // namespace nbx
