/*
 * dgmr -- MI355X (gfx950) rendering hot path of the differentiable Gaussian
 * mixture renderer: the tile splatter and the volumetric ray marcher, forward
 * and backward, as a C ABI.
 *
 * Every pointer in a descriptor is a device pointer unless stated otherwise;
 * camera matrices are host values.  No entry point allocates device memory:
 * outputs, the per-frame cache and scratch come from the caller (the scratch
 * of dgmr_sort_pairs_noalloc is a library-owned array).  All work is
 * enqueued on the given stream; entry points that must know the number of
 * (tile, Gaussian) pairs read one 4-byte counter back from the device (the
 * reference does the same with .item(), splat_forward.cu:141) into a pinned
 * host word and an event that the calling thread takes on its first such
 * call (256 bytes of pinned host memory) and returns to a free list when it
 * ends, for the next thread to reuse.
 *
 * Layouts follow the reference's torch tensors (python_bindings/vol_marcher.cu):
 *   means [n,3], scales [n,3], rotations [n,4] (w,x,y,z), opacity [n],
 *   sh [n,K,3] (K <= 16), matrices [4,4] where tensor row i is glm column i,
 *   images planar [3,H,W].
 */
#ifndef DGMR_H
#define DGMR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* dgmr_stream; /* hipStream_t; NULL = the null stream */

enum dgmr_status {
    DGMR_OK = 0,
    DGMR_ERR_ARGUMENT = 1,   /* bad size, null pointer, unsupported option */
    DGMR_ERR_CAPACITY = 2,   /* pair buffers too small; stats.n_pairs holds the size needed */
    DGMR_ERR_WORKSPACE = 3,  /* workspace too small; query dgmr_*_workspace_size */
    DGMR_ERR_DEVICE = 4      /* a HIP call failed; dgmr_last_error() has the text */
};

enum dgmr_dtype { DGMR_F32 = 0, DGMR_F64 = 1 };

/* DGMR_ACT_NONE: scales, rotations and opacity are used as given (the
 * reference's convention).  DGMR_ACT_LOG_NORM_SIGMOID (splatter only; no
 * counterpart in the reference, which leaves activations to the caller):
 * they are log-scales, unnormalised quaternions and opacity logits, activated
 * inside the preprocess (exp, q/|q|, sigmoid) and differentiated in the
 * backward, so the gradients are w.r.t. the stored values. */
enum dgmr_activation { DGMR_ACT_NONE = 0, DGMR_ACT_LOG_NORM_SIGMOID = 1 };

/* Gaussian parameters as the caller holds them (AoS, contiguous). */
typedef struct dgmr_gaussians {
    int64_t n;
    int32_t dtype;            /* dgmr_dtype; the splatter accepts DGMR_F32 only */
    const void* means;        /* [n,3] */
    const void* scales;       /* [n,3] */
    const void* rotations;    /* [n,4] w,x,y,z */
    const void* opacity;      /* [n]   */
    const void* sh;           /* [n,sh_coeffs,3] */
    int32_t sh_coeffs;        /* K: coefficients stored per Gaussian, 1..16 */
    int32_t sh_degree;        /* active degree 0..3 */
    double scale_multiplier;  /* ForwardData::cov_scale_multiplier (1 in the reference binding) */
    int32_t activation;       /* dgmr_activation: how scales, rotations and opacity are stored */
} dgmr_gaussians;

/* Pinhole camera, host values.  Replaces the .cpu() copies of
 * python_bindings/vol_marcher.cu:16-40. */
typedef struct dgmr_camera {
    double view[16];          /* world_view_transform, tensor layout */
    double proj[16];          /* full_proj_transform, tensor layout */
    double position[3];
    double background[3];
    double tan_fovx, tan_fovy;
    int32_t width, height;
} dgmr_camera;

/* Tile shape and LDS batching.  The reference fixes 16x16 tiles
 * (constants.h:45-46) and 256-Gaussian batches (splat_forward.cu:268-270).
 * cull = 0 bins each Gaussian into its whole 3-sigma tile rectangle, as the
 * reference does (util::getRect, util.h:42; splat_forward.cu:100, 173), so
 * point_list and ranges are the reference's.  cull = 1 (splatter only; no
 * counterpart in the reference) bins it only into the tiles holding a pixel
 * where its alpha can reach 1/255: the blend skips every other (pixel,
 * Gaussian), so images, transmittances and gradients are unchanged while
 * about a fifth fewer pairs are emitted, sorted and blended; point_list and
 * ranges then differ from the reference's.  The marcher and the rasteriser
 * take cull = 0 only. */
typedef struct dgmr_tiling {
    int32_t tile_w, tile_h;   /* 8x8, 16x16 or 32x8 */
    int32_t batch;            /* Gaussians staged in LDS per round: 64, 128 or 256 */
    int32_t cull;             /* 0: the reference's rectangles; 1: tiles reaching alpha 1/255 */
    int32_t slot_order;       /* batched views: 0 = Gaussians take slots in the Morton order of
                                 their centres, 1 = in id order (fewer scattered gradient
                                 writes when a backward returns many gradient groups);
                                 the backward takes the forward's value; other paths ignore it */
} dgmr_tiling;

typedef struct dgmr_stats {
    int64_t n_pairs;          /* (tile, Gaussian) pairs R of this frame */
    int64_t n_visible;        /* Gaussians with at least one tile */
} dgmr_stats;

/* ------------------------------------------------------------ splatter
 * Per-Gaussian cache of the splatter forward (splat_forward.cu:45-64). */
typedef struct dgmr_splat_geometry {
    int32_t* rects;           /* [n,2] pixel extent (ceil(3 sigma)) */
    uint32_t* tiles_touched;  /* [n]   0 = culled */
    float* xy;                /* [n,2] screen centroid */
    float* conic_opacity;     /* [n,4] inverse 2D covariance (xx,xy,yy) + screen opacity */
    float* rgb;               /* [n,3] */
    uint8_t* rgb_clamped;     /* [n,3] 1 where the SH colour was >= 0 */
    float* depths;            /* [n]   */
} dgmr_splat_geometry;

/* Binning cache shared by both renderers (splat_forward.cu:120-240). */
typedef struct dgmr_binning {
    uint32_t* point_offsets;  /* [n] inclusive scan of tiles_touched in Gaussian order */
    uint32_t* point_list;     /* [pair_capacity] Gaussian ids sorted by (tile, depth) */
    uint32_t* ranges;         /* [tiles,2] [first,last) into point_list, 0 for empty tiles */
    int64_t pair_capacity;
} dgmr_binning;

/* Bytes of scratch the splatter / marcher entry points need. */
int dgmr_workspace_size(int64_t n, int32_t width, int32_t height, const dgmr_tiling* tiling, int64_t pair_capacity,
    size_t* bytes);

/* Splatter forward: replaces dgmr::splat_forward (splat_forward.cu:30) /
 * sibr_interfaces::splat (sibr_interfaces.cpp:26-29).  Writes framebuffer
 * [3,H,W] and final transmittance [H,W]; fills geometry and binning. */
int dgmr_splat_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling,
    dgmr_splat_geometry* geom, dgmr_binning* bins, float* framebuffer, float* final_transmittance,
    void* workspace, size_t workspace_bytes, dgmr_stats* stats, dgmr_stream stream);

/* Splatter backward (no counterpart in the reference, which ships no splat
 * backward; its adjoints are composed from grad/math.h and
 * unittests/grad_inria_splatter_render.cu:67-156).  grad_* outputs are
 * overwritten: means [n,3], scales [n,3], rotations [n,4], opacity [n],
 * sh [n,K,3].  screen_grads is [n,9] scratch (xy 2, conic 3, opacity 1, rgb 3). */
int dgmr_splat_backward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling,
    const dgmr_splat_geometry* geom, const dgmr_binning* bins, int64_t n_pairs, const float* framebuffer,
    const float* final_transmittance, const float* grad_framebuffer, float* screen_grads,
    float* grad_means, float* grad_scales, float* grad_rotations, float* grad_opacity, float* grad_sh,
    dgmr_stream stream);

/* ------------------------------------------------------------ marcher
 * Per-Gaussian cache of the marcher forward (vol_marcher_forward.h:28-41);
 * arrays of the scene dtype unless stated. */
typedef struct dgmr_march_geometry {
    int32_t* rects;           /* [n,2] */
    void* rgb;                /* [n,3] */
    uint8_t* rgb_clamped;     /* [n,3] */
    void* depths;             /* [n] */
    void* xy;                 /* [n,2] */
    void* inverse_filtered_cov3d; /* [n,6] xx,xy,xz,yy,yz,zz */
    void* filtered_masses;    /* [n] */
    uint32_t* tiles_touched;  /* [n] */
} dgmr_march_geometry;

typedef struct dgmr_march_options {
    int32_t max_large_steps;  /* 0 = unbounded (the reference); >0 caps large steps per ray */
    int32_t n_frozen;         /* >0: fixed sample sections for gradient checks
                                 (DGMR_TORCH_GRAD_CHECK_CONST_SAMPLES, vol_marcher.h:25) */
    double frozen[8][3];      /* start, end, step of each frozen section */
    int32_t bins_per_chunk;   /* float path: 16 or 32 bins kept in registers per pass (0: 32); results do not depend on it */
    /* ForwardData::debug_render_mode, debug_render_bin, max_depth
     * (vol_marcher.h:44-55; the blends of vol_marcher_forward.cu:443-484),
     * forward only (the backward computes the full mode's gradients):
     * 0 full; 1 bins: one large step, colour (r m + g m + b m + m) * max_depth
     * of bin debug_render_bin (-1, as unsigned, selects the last), T = 0;
     * 2 depth: T multiplies exp(-m) over every bin, colour = the section
     * list's end() / max_depth (+inf unless the list filled up). */
    int32_t render_mode;
    int32_t debug_render_bin;
    double max_depth;         /* <= 0: 20, the reference's default */
    /* Replay store (no counterpart in the reference, whose backward
     * recomputes every bin): device memory of dgmr_march_replay_size bytes,
     * 256-byte aligned, or null.  The full-mode forward keeps in it, for the
     * first replay_steps large steps of every tile, each pixel's bins and
     * bin borders; a backward given the same store (and the same scene,
     * camera, tiling and options, with nothing else written to it between
     * the two calls) reads them back instead of recomputing them, for the
     * tiles whose rays took at most replay_steps large steps, and recomputes
     * the rest.  Results are the same either way. */
    void* replay;
    size_t replay_bytes;
    int32_t replay_steps;
} dgmr_march_options;

/* Bytes of a replay store keeping replay_steps large steps per tile
 * (dtype: dgmr_dtype of the scene).  About 2.5 KB (float) per pixel and
 * kept step. */
int dgmr_march_replay_size(int32_t width, int32_t height, const dgmr_tiling* tiling, int32_t replay_steps, int32_t dtype, size_t* bytes);

/* Marcher forward: replaces dgmr::vol_marcher::forward<T>
 * (vol_marcher_forward.cu:42-43).  framebuffer [3,H,W], remaining
 * transparency [H,W], both of the scene dtype. */
int dgmr_march_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling,
    const dgmr_march_options* opt, dgmr_march_geometry* geom, dgmr_binning* bins, void* framebuffer,
    void* remaining_transparency, void* workspace, size_t workspace_bytes, dgmr_stats* stats, dgmr_stream stream);

/* Marcher backward: replaces dgmr::vol_marcher::backward<T>
 * (vol_marcher_backward.cu:53-54).  render_grads is [n,13] scratch of the
 * scene dtype (mass 1, rgb 3, means 3, inverse covariance 6).  grad_sh is
 * [n,K,3]; SH coefficients beyond K read as 1 in this pass, as the reference
 * binding pads them (python_bindings/vol_marcher.cu:177). */
int dgmr_march_backward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_tiling* tiling,
    const dgmr_march_options* opt, const dgmr_march_geometry* geom, const dgmr_binning* bins, int64_t n_pairs,
    const void* framebuffer, const void* remaining_transparency, const void* grad_framebuffer, void* render_grads,
    void* grad_means, void* grad_scales, void* grad_rotations, void* grad_opacity, void* grad_sh, dgmr_stream stream);

/* ------------------------------------------------------------ batched views
 * Many cameras of one scene (uneven resolutions) rendered together: one
 * preprocess launch over all (view, Gaussian) pairs, one binning pass over the
 * global tile range of all views, and one persistent blend launch whose
 * workgroups pull tiles of every view from a device queue.  No counterpart in
 * the reference (one camera per call there, vol_marcher.py:19-65).
 * The workspace carries the per-view geometry and binning from the forward to
 * the backward: keep it untouched in between. */
typedef struct dgmr_batch_view {
    dgmr_camera cam;
    float* framebuffer;           /* [3,H,W] output of the forward */
    float* final_transmittance;   /* [H,W]   output of the forward */
    const float* grad_framebuffer;/* [3,H,W] input of the backward */
    /* Fused L1 loss (no counterpart in the reference): l1_target non-NULL
     * makes the backward use dL/dimage = sign(framebuffer - l1_target) *
     * l1_scale at every pixel, bit for bit what dgmr_l1_loss_backward writes
     * with scale l1_scale and dL/dloss = 1, and ignore grad_framebuffer. */
    const float* l1_target;       /* [3,H,W] or NULL */
    float l1_scale;               /* 1 / (3 H W) for the mean */
    int32_t reserved;
} dgmr_batch_view;

int dgmr_batch_workspace_size(int64_t n, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling,
    int64_t pair_capacity, size_t* bytes);

int dgmr_splat_batch_forward(const dgmr_gaussians* g, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling,
    void* workspace, size_t workspace_bytes, int64_t pair_capacity, dgmr_stats* stats, dgmr_stream stream);

/* Gradient groups: view_group NULL sums the gradients of all views into
 * grad_* (one block, n_groups ignored).  Otherwise view_group (host, n_views
 * entries) gives each view a group in [0, n_groups) or -1 (no gradient); a
 * group's views are consecutive, and grad_* hold n_groups consecutive blocks
 * ([n_groups][n,3] means, ...), each the sum over its group's views (zero for
 * a group without views).  A view whose grad_framebuffer and l1_target are
 * both NULL contributes nothing, as a zero dL/dimage would (group -1 views
 * need neither).  Outputs are overwritten. */
int dgmr_splat_batch_backward(const dgmr_gaussians* g, const dgmr_batch_view* views, int32_t n_views, const dgmr_tiling* tiling,
    void* workspace, size_t workspace_bytes, int64_t pair_capacity, int64_t n_pairs, const int32_t* view_group, int32_t n_groups,
    float* grad_means, float* grad_scales, float* grad_rotations, float* grad_opacity, float* grad_sh, dgmr_stream stream);

/* ------------------------------------------------------------ optimiser step
 * Adam over float32 parameter tensors, the step that follows the backward in
 * a 3DGS training loop (means, log-scales, quaternions, opacity logits, SH).
 * No counterpart in the reference, which returns gradients only
 * (python_bindings/vol_marcher.py:48-65).  Per element, as torch.optim.Adam
 * (L2 weight decay added to the gradient):
 *   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
 *   p -= lr / (1-b1^step) * m / (sqrt(v) / sqrt(1-b2^step) + eps). */
typedef struct dgmr_adam_tensor {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t numel;
    float lr;
} dgmr_adam_tensor;

int dgmr_adam_step(const dgmr_adam_tensor* tensors, int32_t n_tensors, double beta1, double beta2, double eps, double weight_decay,
    int64_t step, dgmr_stream stream);

/* ------------------------------------------------------------ stage timing
 * When enabled, every entry point brackets each kernel stage with hipEvents on
 * its stream (no extra synchronisation).  dgmr_profile_read waits for the
 * recorded events and returns per-stage totals; names are static strings.
 * No counterpart in the reference (its timers are commented out,
 * sibr_interfaces.cpp:33-43). */
typedef struct dgmr_stage_time {
    const char* name;
    double total_ms;
    int64_t launches;
    double bytes;             /* algorithmic bytes summed over launches (0 if not modelled) */
} dgmr_stage_time;

int dgmr_profile_enable(int on);
int dgmr_profile_read(dgmr_stage_time* out, int32_t max_stages, int32_t* n_stages);

/* ------------------------------------------------------------ rasteriser
 * dgmr::vol_raster_forward(VolRasterForwardData&) (vol_raster_forward.h:68,
 * vol_raster_forward.cu:36-520): forward only, as the reference.  16x16
 * tiles; the binning (geom, bins, workspace) is the splatter's, sized with
 * dgmr_workspace_size and a 16x16 tiling.  inv_cov [N,6] and mass [N]
 * (device) receive the inverse filtered covariance and the filtered mass of
 * each visible Gaussian.  render_mode: 0 full, 1 bins (background only, as
 * the reference's unfinished mode), 2 depth (end of bin debug_bin / max_depth).
 * With no Gaussian binned the framebuffer is left untouched, as the reference
 * returns before rendering. */
typedef struct {
    int32_t render_mode;
    int32_t debug_bin;  /* -1 selects the last bin */
    float max_depth;    /* depth mode scale (the reference's default 20) */
} dgmr_raster_options;

int dgmr_raster_forward(const dgmr_gaussians* g, const dgmr_camera* cam, const dgmr_raster_options* opt, dgmr_splat_geometry* geom,
    dgmr_binning* bins, float* inv_cov, float* mass, float* framebuffer, void* workspace, size_t workspace_bytes, dgmr_stats* stats,
    dgmr_stream stream);

/* ------------------------------------------------------------ L1 loss
 * No counterpart in the reference, which leaves the loss to its caller: the
 * photometric L1 loss of a training step, L = scale * sum |image - target|
 * (scale = 1/n for the mean), fused into one pass each way.  The forward
 * writes dgmr_l1_partials_size(n) scaled partial sums to partials (device),
 * whose sum is L; the backward writes grad = sign(image - target) *
 * scale * grad_loss[0] (grad_loss on the device).  All arrays float32,
 * 16-byte aligned. */
int dgmr_l1_partials_size(int64_t n, int32_t* count);
int dgmr_l1_loss_forward(const float* image, const float* target, int64_t n, float scale, float* partials, dgmr_stream stream);
int dgmr_l1_loss_backward(const float* image, const float* target, int64_t n, float scale, const float* grad_loss, float* grad,
    dgmr_stream stream);

/* The backward over many images in one launch (the views of a batch): item k
 * writes items[k].grad = sign(image - target) * scale * grad_loss[0], as
 * dgmr_l1_loss_backward does for one image.  The descriptors are host memory;
 * the arrays they point to are device memory, float32, 16-byte aligned. */
typedef struct dgmr_l1_item {
    const float* image;
    const float* target;
    float* grad;
    const float* grad_loss; /* device scalar dL/dloss */
    int64_t n;
    float scale;
    int32_t reserved;
} dgmr_l1_item;

int dgmr_l1_loss_backward_many(const dgmr_l1_item* items, int32_t n_items, dgmr_stream stream);

/* ------------------------------------------------------------ sort
 * Stable LSD radix sort of n (uint32 key, uint32 value) pairs on key bits
 * [begin_bit, end_bit), the primitive both binnings use (replaces
 * cub::DeviceRadixSort::SortPairs, splat_forward.cu:193-211 and
 * vol_marcher_forward.cu:247-263).  keys/vals (device, n each) are sorted in
 * place; keys_tmp/vals_tmp (device, n each) are the other half of the
 * ping-pong; ws (device) holds at least dgmr_sort_workspace_size(n) bytes.
 * Bits outside [begin_bit, end_bit) do not take part in the order. */
int dgmr_sort_workspace_size(int64_t n, size_t* bytes);
int dgmr_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int32_t begin_bit, int32_t end_bit,
    void* ws, size_t ws_bytes, dgmr_stream stream);

/* ------------------------------------------------------------ scan
 * Prefix sum of n uint32 values (device), the primitive under every binning
 * offset (replaces cub::DeviceScan::InclusiveSum, splat_forward.cu:127-137,
 * vol_marcher_forward.cu:185-189, vol_raster_forward.cu:146-149).  inclusive
 * != 0: out[i] = in[0] + ... + in[i]; 0: out[i] = in[0] + ... + in[i-1].
 * Sums wrap modulo 2^32.  total (device, may be null) receives the sum of all
 * n.  in may equal out.  ws (device) holds at least dgmr_scan_workspace_size(n)
 * bytes; its contents need no initialisation. */
int dgmr_scan_workspace_size(int64_t n, size_t* bytes);
int dgmr_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, int32_t inclusive, void* ws, size_t ws_bytes,
    dgmr_stream stream);

/* No counterpart in the reference: dgmr_sort_pairs for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same sort, with the scratch that
 * dgmr_sort_pairs takes as ws kept in a library-owned device array of
 * 16 MiB + 64 KiB per GPU (allocated with the library's code object and held
 * until the process ends); the caller still passes the keys, values and the
 * ping-pong halves.  n above DGMR_SORT_NOALLOC_MAX_ITEMS (2^25) is rejected
 * with DGMR_ERR_ARGUMENT before any device work.  The entry point calls no
 * allocating routine.  Every call on one GPU uses that GPU's one array:
 * calls in stream order are safe, calls that overlap on one GPU (other
 * streams, other threads) race. */
#define DGMR_SORT_NOALLOC_MAX_ITEMS 33554432
int dgmr_sort_pairs_noalloc(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, int64_t n, int32_t begin_bit,
    int32_t end_bit, dgmr_stream stream);

/* ------------------------------------------------------------ self-test
 * Runs the blend backward's wave reductions on one wave: in holds 9 x 64
 * floats (device), out receives 128 (device): out[8m] = sum over lanes of
 * in[q*64 + lane] for q = {0,4,2,6,1,5,3,7}[m], out[64 + 63] = sum of in[512..575].
 * Diagnostic only; no counterpart in the reference. */
int dgmr_selftest_wave_sum(const float* in, float* out, dgmr_stream stream);

/* ------------------------------------------------------------ misc */
const char* dgmr_last_error(void);
const char* dgmr_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DGMR_H */
