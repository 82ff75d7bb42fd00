/*
 * pxl.h -- C ABI of libpxl, the MI355X (gfx950) kernels of the voxel-grid
 * render-and-optimise path.
 *
 * The reference (DanJbk/Plenoxels) has no FFI: its boundary is the Python call signatures
 * of src/ray_sampling.py and src/grid_functions.py as called from scripts/train.py:130-157
 * (training, with autograd backward) and src/visualization.py:125-146 (forward-only render).
 * Each entry point below names the reference function(s) whose work it replaces.
 *
 * Conventions
 *   - Plain C: pointers are device pointers (hipMalloc'd or torch-owned), sizes are int64.
 *   - The library never allocates device memory; the caller owns every buffer.
 *   - Work is enqueued asynchronously on `stream` (a hipStream_t; NULL = default stream).
 *   - Return 0 on success or a negative PXL_E_* code; pxl_last_error() then describes it
 *     (thread-local).  Argument validation happens before any HIP call.
 *   - A call over zero rays / samples / elements succeeds and launches nothing, whatever the
 *     pointers of its per-ray or per-element buffers (torch gives an empty tensor a NULL data
 *     pointer); descriptors, the grid and grid-shaped buffers are still checked.
 *   - Grid layout (src/grid_functions.py:214-215): row-major (X,Y,Z,4), channel innermost,
 *     channels [R,G,B,alpha]; linear cell index L = (i*Y + j)*Z + k.  Gradients are always
 *     FP32 (X,Y,Z,4) and are ACCUMULATED into (the caller zero-fills when it wants d/dgrid).
 *   - Gradient accumulation uses float atomics: summation order, hence the last bits of
 *     a gradient, may differ run to run.
 */
#ifndef PXL_H
#define PXL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXL_ABI_VERSION 3

/* return codes */
#define PXL_OK 0
#define PXL_E_INVALID (-1)   /* bad argument (null pointer, bad enum, negative size) */
#define PXL_E_DTYPE (-2)     /* unsupported storage dtype / channel count */
#define PXL_E_SIZE (-3)      /* a size exceeds what the kernels index (e.g. >= 2^31 cells) */
#define PXL_E_HIP (-4)       /* a HIP runtime call or kernel launch failed */
#define PXL_E_NO_DEVICE (-5) /* no HIP device available */

/* grid storage dtypes */
#define PXL_DTYPE_F32 0
#define PXL_DTYPE_BF16 1
#define PXL_DTYPE_F16 2

/* sample-to-grid lookup */
#define PXL_INTERP_NEAREST 0   /* src/grid_functions.py:103-114 */
#define PXL_INTERP_TRILINEAR 1 /* src/grid_functions.py:7-44, :154-170, :220-246 */

/* pxl_march.flags */
#define PXL_FLAG_EXACT_DIV 1 /* n = (p - min)/pd (ATen CPU); default n = (p - min)*inv_pd (ATen GPU) */

/* A dense voxel grid. */
typedef struct pxl_grid {
  const void* data;          /* (X,Y,Z,4) cells in `dtype` */
  int64_t dims[3];           /* X, Y, Z */
  int32_t channels;          /* must be 4 */
  int32_t dtype;             /* PXL_DTYPE_* */
  const uint32_t* occupancy; /* optional occupancy bits in 8x8x8 tiles (pxl_occupancy_words
                                words): tile (x>>3, y>>3, z>>3) is the 16 words from
                                ((tx*TY + ty)*TZ + tz)*16, TY = ceil(Y/8), TZ = ceil(Z/8), and
                                cell (x,y,z) is bit (y&3)*8 + (z&7) of its word (x&7)*2 +
                                ((y>>2)&1); set = occupied.  Nearest skips a sample whose
                                nearest cell is unoccupied; trilinear skips one whose floor
                                cell is unoccupied (see pxl_occupancy_dilate).  NULL = every
                                cell occupied. */
} pxl_grid;

/* How samples are placed along each ray and mapped to grid-index space
 * (src/ray_sampling.py:161-167 and :12-13):
 *   t_k = delta*k + near, k = 1..n_samples;   p = o + d*t_k;   n = (p - grid_min) * inv_pd
 * every step rounded to float32 separately, as the reference's torch ops are. */
typedef struct pxl_march {
  int32_t n_samples;  /* S >= 0 (<= 2^24) */
  int32_t interp;     /* PXL_INTERP_* */
  int32_t flags;      /* PXL_FLAG_* */
  int32_t reserved;   /* 0 */
  float delta;        /* delta_step */
  float near;         /* 0 reproduces the reference */
  float t_stop;       /* early termination: drop samples whose transmittance < t_stop; 0 = off */
  float pd;           /* points_distance as float32 */
  float inv_pd;       /* float32(1/float32(pd)) */
  float grid_min[3];  /* grid_indices.min(0) (src/ray_sampling.py:13) */
} pxl_march;

/* pxl_rays.flags */
#define PXL_RAYS_CPU_NORM 1 /* camera-form directions normalised in the summation order of torch's
                               CPU norm, sqrt(fma(z,z,fma(y,y,x*x))); default: the order of its
                               GPU norm, sqrt((x*x + z*z) + y*y) -- the reference's default device
                               is 'cuda' (src/ray_sampling.py:195,218,262) */

/* A batch of rays, in one of two forms.
 * Explicit: origins/dirs (n_rays,3) f32.
 * Camera (origins == NULL): ray r comes from camera cam_ids[r] at uniforms uv[r] = (u,v) in
 * [0,1], exactly as src/ray_sampling.py:212-264 derives a direction and a target pixel from
 * a c2w matrix and camera_angle_x; its origin is c2w[cam][:3,3] (src/ray_sampling.py:159).
 * Every op is rounded as torch rounds it; the two 3-norms follow `flags`. */
typedef struct pxl_rays {
  int64_t n_rays;
  const float* origins;   /* (n,3) or NULL for camera form */
  const float* dirs;      /* (n,3) */
  const int32_t* cam_ids; /* (n,)  camera form */
  const float* uv;        /* (n,2) camera form */
  const float* c2w;       /* (n_cams,4,4) row-major camera-to-world matrices */
  int64_t n_cams;
  float camera_angle_x;
  int32_t flags;          /* PXL_RAYS_* */
  const float* imgs;      /* (n_cams,H,W,4) target images or NULL */
  int64_t img_h, img_w;
} pxl_rays;

int pxl_abi_version(void);
const char* pxl_last_error(void);

/* Floats of per-ray state pxl_render_fwd saves for pxl_render_bwd. */
int64_t pxl_aux_floats(int32_t n_samples);

/* Fused forward render: sample placement, gather (clip to [0,1], out-of-grid samples
 * zeroed), front-to-back compositing.  out_rgba (n,4) f32; aux (n, pxl_aux_floats(S)) f32
 * or NULL when no backward will follow.
 * Replaces, for one batch: src/ray_sampling.py:161-167 (sample_camera_rays_batched),
 * :12-13 (normalize_samples_for_indecies), src/grid_functions.py:103-114
 * (get_nearest_voxels) or :154-170 + :7-44 (trilinear), scripts/train.py:146-147 (clip,
 * mask), src/ray_sampling.py:172-192 (compute_alpha_weighted_pixels). */
int pxl_render_fwd(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   float* out_rgba, float* aux, void* stream);

/* Fused backward of pxl_render_fwd: accumulates dL/dgrid into grad_grid (X,Y,Z,4) f32
 * given dL/dout (n,4).  Replaces autograd's replay of the chain above
 * (scripts/train.py:181, the backward of :142-153). */
int pxl_render_bwd(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   const float* grad_rgba, const float* aux, float* grad_grid, void* stream);

/* Loss terms of one training step (scripts/train.py:150-177). */
typedef struct pxl_loss {
  const float* targets; /* (n,4) f32, or NULL in camera form with imgs set (targets gathered
                           from imgs as src/ray_sampling.py:238-248 does) */
  float grad_scale;     /* d loss / d out = grad_scale * (out - target); 2/(4*N_rays) is
                           mse_loss's mean (scripts/train.py:156) */
  float beta_scale;     /* beta / N_samples: weight of the per-sample term
                           f(a) = log(a + 1e-4) - log(1 - a + 1e-4) averaged over all samples
                           (scripts/train.py:170-174); 0 disables it */
  float* ray_sq_err;    /* (n,) or NULL: sum over channels of (out - target)^2 */
  float* ray_beta;      /* (n,) or NULL: sum of f(a) - f(0) over the ray's processed samples
                           (every other sample contributes f(0)) */
} pxl_loss;

/* One training step's render + loss gradient + backward in one pass: out = render(rays);
 * accumulates d loss / d grid into grad_grid.  out_rgba may be NULL.  n_samples <=
 * PXL_TRAIN_MAX_SAMPLES: up to 512 samples every kept sample's value stays in registers
 * between the forward and the reverse sweep; above that each ray keeps its per-group state
 * (64 kept samples a group) in LDS and re-gathers each group on the reverse sweep.
 * Replaces scripts/train.py:130-177 + :180-181 minus the optimiser and the TV term. */
#define PXL_TRAIN_MAX_SAMPLES 16384
int pxl_train_step(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                   const pxl_loss* loss, float* out_rgba, float* grad_grid, void* stream);

/* A step's loss value from its per-ray squared errors: out[0] = (float)(scale * sum of
 * v[0, n)), summed in FP64 in a fixed order (the same bits on every call), in one launch.
 * With v = pxl_loss.ray_sq_err and scale = 1 / (4 * N_rays) this is mse_loss's mean
 * (scripts/train.py:156), the value the reference reads every step (:159).  v and work
 * 16-byte aligned; work: pxl_scaled_sum_work_bytes() bytes, zeroed before its first use and
 * left ready for the next call by every call (calls sharing one work buffer must not overlap:
 * keep them on one stream). */
int pxl_scaled_sum(const float* v, int64_t n, double scale, float* out, void* work, void* stream);
int64_t pxl_scaled_sum_work_bytes(void);

/* pxl_scaled_sum without the caller's work buffer: the same kernel and the same bits, with its
 * partials and ticket in a scratch array the library keeps in each GPU's memory
 * (pxl_scaled_sum_noalloc_scratch_bytes() = 1 KiB per GPU, held until the process ends).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc -- and it
 * allocates nothing itself.  Any n >= 0 fits the scratch (its size does not depend on n).  The
 * first call on a device looks the scratch up and clears it with a kernel on `stream`; every
 * call leaves it ready for the next.  All calls on one GPU share the one scratch: calls that
 * overlap on one GPU (on different streams, or from different threads without ordering) race. */
int pxl_scaled_sum_noalloc(const float* v, int64_t n, double scale, float* out, void* stream);
int64_t pxl_scaled_sum_noalloc_scratch_bytes(void);

/* pxl_train_step with the trilinear gradient accumulated into corner blocks instead of
 * grad_grid: blocks is (X*Y*Z, 16) f32 (pxl_blocks_floats), zero on entry: one 64-byte entry
 * per cell E = (x, y, z), whose slot s = 2bx + by (floats 4s + c) receives channel c of corner
 * (bx, by, bz) -- b = 0 the ceil index, b = 1 the floor index, the corner order of
 * src/grid_functions.py:238-243 -- of every kept sample whose floor cell is (x, y, zf) and
 * whose corner z (zf for bz = 1, the wrapped ceil of zf for bz = 0) is z, clip mask applied.
 * A sample's 32 atomics then fall in two 64-byte entries (two memory requests instead of about
 * five for the cells themselves).  pxl_blocks_reduce turns the blocks into the (X,Y,Z,4)
 * gradient.  Several calls may accumulate into the same blocks before one reduce.  interp must
 * be trilinear; at most 2^30 - 1 cells.  Same replacement as pxl_train_step. */
int pxl_train_step_blocks(const pxl_grid* grid, const pxl_march* march, const pxl_rays* rays,
                          const pxl_loss* loss, float* out_rgba, float* blocks, void* stream);

/* Floats of the corner blocks of a grid (X*Y*Z*16), or -1 for invalid dims. */
int64_t pxl_blocks_floats(const pxl_grid* grid);

/* pxl_blocks_reduce modes */
#define PXL_REDUCE_OVERWRITE 0            /* every cell of grad_grid written */
#define PXL_REDUCE_ACCUMULATE 1           /* every cell added to */
#define PXL_REDUCE_OVERWRITE_REACHABLE 2  /* only the cells some written entry reaches (with
                                             grid->occupancy: those with a source entry of a set
                                             bit); the others are not touched.  For a gradient
                                             buffer that holds zeros outside that reach -- zeroed
                                             once and used with one occupancy -- the result equals
                                             PXL_REDUCE_OVERWRITE without its dense store pass. */

/* grad_grid (X,Y,Z,4) f32 = (accumulate: +=) the sum, for every cell (x, y, z), of slot
 * 2bx + by of entry (x - 1 + bx, y - 1 + by, z) over bx, by in {0, 1} (corner (bx, by) of row
 * (x, y) names cell (x + 1 - bx, y + 1 - by), the ceil of the last index wrapping to 0 as
 * src/grid_functions.py:168 does).  When grid->occupancy is set (pass the occupancy the
 * training steps used) only entries that such steps can write are read: those of cells whose
 * bit, or whose z - 1 neighbour's bit (wrapped), is set; grid->data is not read.  clear:
 * blocks to zero for the next training step -- == blocks: these blocks, in a second pass after
 * the sum; another buffer of the same size: its entries, in the same pass (double buffering:
 * the next step accumulates into `clear` while this one's blocks wait for their own clearing
 * pass); NULL: none.  Only those same entries are zeroed when occupancy is set, so keep the
 * occupancy of a block buffer fixed or zero it whole when it changes. */
int pxl_blocks_reduce(const pxl_grid* grid, float* blocks, float* grad_grid, int32_t accumulate,
                      float* clear, void* stream);  /* accumulate: PXL_REDUCE_* */

/* Zero the entries of corner blocks that a training step with grid->occupancy can write (every
 * entry when occupancy is NULL) -- the clearing pxl_blocks_reduce does with clear == blocks,
 * as a call of its own, e.g. on a second stream while the other buffer of a pair is in use.
 * grid->data is not read. */
int pxl_blocks_clear(const pxl_grid* grid, float* blocks, void* stream);

/* Ray generation only (src/ray_sampling.py:195-264 on per-ray camera/uniforms):
 * origins (n,3), unit dirs (n,3), targets (n,4) (NULL to skip; needs imgs). */
int pxl_camera_rays(const pxl_rays* rays, float* origins, float* dirs, float* targets,
                    void* stream);

/* Cell occupancy for trilinear skipping: the out bit of cell (i,j,k) = OR of the in bits of
 * its 8 corners (i+a, j+b, k+c), a,b,c in {0,1}, wrapped.  Both in the tile layout of
 * pxl_grid.occupancy (pxl_occupancy_words words); out must not alias in. */
int pxl_occupancy_dilate(const uint32_t* occ_in, int64_t X, int64_t Y, int64_t Z,
                         uint32_t* occ_out, void* stream);

/* Occupancy from the grid's alpha channel: the bit of cell L in bits (pxl_occupancy_words
 * words, the layout of pxl_grid.occupancy) set when raw alpha >= 0 at cell L, or, with
 * dilate, at any of its 8
 * corners (the cells a trilinear sample with floor cell L reads).  Samples that this skips are
 * exactly zero in the reference chain (clamp gives alpha 0 and a zero gradient mask), so
 * passing it to the render kernels changes no finite result; NaN alpha counts as empty.
 * grid->occupancy is not read. */
int pxl_alpha_occupancy(const pxl_grid* grid, int32_t dilate, uint32_t* bits, void* stream);

/* Pack a (X,Y,Z) occupancy grid of bytes (nonzero = occupied) into the tile layout of
 * pxl_grid.occupancy.  out: pxl_occupancy_words(X, Y, Z) words (padding bits 0). */
int pxl_pack_occupancy(const uint8_t* occ, int64_t X, int64_t Y, int64_t Z, uint32_t* out,
                       void* stream);

/* Words of the occupancy bits of an (X,Y,Z) grid: 16 per 8x8x8 tile, dims rounded up to a
 * multiple of 8; -1 for invalid dims. */
int64_t pxl_occupancy_words(int64_t X, int64_t Y, int64_t Z);

/* dst[0..n) = value (HBM-rate fill, e.g. the per-step gradient zero-fill). */
int pxl_fill_f32(float* dst, int64_t n, float value, void* stream);

/* The gradient zero-fill of a training step with grid->occupancy: grad (X,Y,Z,4) f32 = value
 * at every cell the step's scatter can write -- nearest: each occupied cell; trilinear (corners
 * one cell above the floor cell, wrapped): every cell of each 8x8x8 tile whose own bits or
 * those of a lower neighbour tile (x-1, y-1, z-1, wrapped) are set.  Other cells are not
 * touched: a gradient buffer zeroed once and used with one occupancy then holds exactly what a
 * dense zero-fill would give it.  grid->data is not read; grid->occupancy is required. */
int pxl_fill_reachable_f32(const pxl_grid* grid, int32_t interp, float* grad, float value, void* stream);

/* ---------------------------------------------------------------------------------------
 * Per-function kernels behind the reference's individual call signatures.  Unless stated
 * otherwise: float32, contiguous, device pointers; N = number of points; grids (X,Y,Z,4).
 * ------------------------------------------------------------------------------------- */

/* src/ray_sampling.py:161-167: out[(r*S + k-1)] = o + d_r * (delta*k + near), k = 1..S, with
 * o = origins[r / rays_per_origin] (the reference repeats each camera position over its
 * rays).  out: (n_rays*S, 3). */
int pxl_sample_points(const float* origins, int64_t rays_per_origin, const float* dirs,
                      int64_t n_rays, int32_t n_samples, float delta, float near, float* out,
                      void* stream);

/* src/ray_sampling.py:13 grid_indices.min(0)[0]: out3 (3 floats, device) = column minima of
 * coords (n,3). */
int pxl_coord_min(const float* coords, int64_t n, float* out3, void* stream);

/* src/ray_sampling.py:12-13: out = (p - min3) * inv_pd (or / pd with PXL_FLAG_EXACT_DIV);
 * min3 is a device pointer (3 floats). */
int pxl_normalize_points(const float* p, int64_t n, const float* min3, float pd, float inv_pd,
                         int32_t flags, float* out, void* stream);

/* src/grid_functions.py:47-63 (find_out_of_bound, True = inside): idx (N,3) int64 when
 * is_float == 0, float32 otherwise; mask (N,) bytes. */
int pxl_in_bounds(const void* idx, int32_t is_float, int64_t n, const int64_t dims[3],
                  uint8_t* mask, void* stream);

/* src/grid_functions.py:66-79 (fix_out_of_bounds): idx (N,3) int64 wrapped in place,
 * result in [0,dim) (Python modulo). */
int pxl_wrap_indices(int64_t* idx, int64_t n, const int64_t dims[3], void* stream);

/* src/grid_functions.py:103-114 (get_nearest_voxels): idx = round-half-even(n), mask =
 * in-bounds, vals = grid[idx wrapped] (every point, as the reference gathers before the
 * caller applies the mask).  vals (N,4), mask (N,) bytes.  grid: f32. */
int pxl_nearest_gather(const float* n, int64_t N, const pxl_grid* grid, float* vals,
                       uint8_t* mask, void* stream);
/* backward of the gather: grad_grid[idx wrapped] += grad_vals (N,4) */
int pxl_nearest_scatter(const float* n, int64_t N, const pxl_grid* grid, const float* grad_vals,
                        float* grad_grid, void* stream);

/* src/grid_functions.py:220-246 (get_grid_points_indices): corners (N,8,3) int64,
 * x{ceil,floor} x y{ceil,floor} x z{ceil,floor}. */
int pxl_corner_indices(const float* n, int64_t N, int64_t* corners, void* stream);

/* src/grid_functions.py:154-170 (collect_cell_information_via_indices): out (N*8,4) = grid at
 * the wrapped corners, mask (N,) = float in-bounds test.  grid: f32. */
int pxl_collect_cells(const float* n, int64_t N, const pxl_grid* grid, float* out, uint8_t* mask,
                      void* stream);
int pxl_collect_cells_bwd(const float* n, int64_t N, const pxl_grid* grid, const float* grad_out,
                          float* grad_grid, void* stream);

/* src/grid_functions.py:7-44 (trilinear_interpolation): out (N,4) from corners (N,8,3) int64
 * (pre-wrapped, as the reference requires; out-of-range corners are wrapped here instead of
 * raising) with signed frac weights, in the reference's lerp order.  grid: f32. */
int pxl_trilinear(const float* n, const int64_t* corners, int64_t N, const pxl_grid* grid,
                  float* out, void* stream);
int pxl_trilinear_bwd(const float* n, const int64_t* corners, int64_t N, const pxl_grid* grid,
                      const float* grad_out, float* grad_grid, void* stream);

/* src/ray_sampling.py:172-192 (compute_alpha_weighted_pixels): samples (R,S,4) -> out (R,4);
 * backward writes grad_samples (R,S,4) (overwritten, not accumulated).  S <= 65536. */
int pxl_composite_fwd(const float* samples, int64_t R, int32_t S, float* out, void* stream);
int pxl_composite_bwd(const float* samples, int64_t R, int32_t S, const float* grad_out,
                      float* grad_samples, void* stream);

/* src/grid_functions.py:198-212 (generate_grid's coordinates): coords (X,Y,Z,3) f32 with
 * cell (i,j,k) at ((i,j,k) - ceil(S/2) + 1) * pd; mesh (X,Y,Z,3) int64 or NULL. */
int pxl_grid_coords(int64_t X, int64_t Y, int64_t Z, float pd, float* coords, int64_t* mesh,
                    void* stream);

/* src/grid_functions.py:173-181 (average_pool3d_grid): avg_pool3d with kernel k, stride s, no
 * padding over the (X,Y,Z,4) layout directly; out (Xo,Yo,Zo,4), Xo = (X-k)/s + 1.  Computed as
 * the mean over k along z, y, then x (the window is separable); workspace: caller's scratch of
 * pxl_avg_pool3d_workspace_floats(X,Y,Z,k,s) floats. */
int64_t pxl_avg_pool3d_workspace_floats(int64_t X, int64_t Y, int64_t Z, int32_t k, int32_t s);
int pxl_avg_pool3d(const float* grid, int64_t X, int64_t Y, int64_t Z, int32_t k, int32_t s,
                   float* out, float* workspace, void* stream);
/* backward: grad_grid (X,Y,Z,4) = (accumulate: +=) sum over windows containing a cell of
 * grad_out / k^3 (the adjoint passes x, y, z; every cell is written, 0 outside all windows);
 * workspace as for the forward. */
int pxl_avg_pool3d_bwd(const float* grad_out, int64_t X, int64_t Y, int64_t Z, int32_t k,
                       int32_t s, float* grad_grid, int32_t accumulate,
                       float* workspace, void* stream);

/* scripts/train.py:44-65 (tv_loss): sum_sq (1 double, device, accumulated) += sum over the
 * three axes of squared neighbour differences of grid (X,Y,Z,4). */
int pxl_tv_sumsq(const float* grid, int64_t X, int64_t Y, int64_t Z, double* sum_sq, void* stream);
/* d(weight * sqrt(sum_sq))/dgrid accumulated into grad_grid; sum_sq read on the device. */
int pxl_tv_bwd(const float* grid, int64_t X, int64_t Y, int64_t Z, const double* sum_sq,
               float weight, float* grad_grid, void* stream);

/* torch.optim.Adam step (scripts/train.py:89,182; defaults betas (0.9,0.999), eps 1e-8, no
 * weight decay) fused with the |grad| accumulation of scripts/train.py:184:
 *   m = lerp(m, g, 1-b1); v = b2 v + (1-b2) g^2;
 *   p -= lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps);  grad_abs += |g|
 * The scalar coefficients are formed in double and rounded to float32, as torch forms them
 * in Python.  param/exp_avg/exp_avg_sq/grad FP32 (n).  grad_abs may be NULL.  half_out (n) of
 * half_dtype (PXL_DTYPE_BF16 / PXL_DTYPE_F16, round to nearest even) receives the updated
 * parameter for a reduced-precision render copy, or NULL. */
int pxl_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                  float* grad_abs, int64_t n, double lr, double beta1, double beta2, double eps,
                  int64_t step, void* half_out, int32_t half_dtype, void* stream);

/* pxl_tv_bwd followed by pxl_adam_step in one pass over an (X,Y,Z,4) grid (scripts/train.py:
 * 163-168 and :180-184 for a step with the TV term): the total gradient grad + d(tv_weight *
 * sqrt(tv_sum_sq))/dparam is formed in registers, exactly as pxl_tv_bwd would add it into grad
 * (grad itself is not written), and drives the Adam update, the |grad| accumulation and the
 * half copy; the result equals the two separate calls bit for bit.  The updated parameter is
 * written to param_out, which must be a different buffer from param_in: the TV stencil reads
 * each cell's six neighbours from param_in.  grad is not written unless zero_grad is nonzero:
 * then every grad element read as nonzero is set to 0 after use, so grad is all zero on
 * return and the next step's scatter needs no zero-fill (a step's gradient covers few cells).
 * Arrays 16-byte aligned (half_out 8-byte). */
int pxl_adam_tv_step(const float* param_in, float* param_out, float* grad, float* exp_avg,
                     float* exp_avg_sq, float* grad_abs, int64_t X, int64_t Y, int64_t Z,
                     const double* tv_sum_sq, float tv_weight, double lr, double beta1, double beta2,
                     double eps, int64_t step, void* half_out, int32_t half_dtype, int32_t zero_grad,
                     void* stream);

/* Host-side traffic accounting (no counterpart in the reference; ALL POINTERS ARE HOST
 * POINTERS, including grid->occupancy, and grid->data is not read): counts[r] = number of
 * distinct grid cells the processed samples of ray r read -- nearest cell, or the union of
 * the 8 trilinear corners, of every in-grid, non-skipped sample k <= n_proc[r] (n_proc may
 * be NULL = all samples; pxl_render_fwd stores it in aux[r][0]).  Multithreaded. */
int pxl_count_cells_host(const pxl_grid* grid, const pxl_march* march, const float* origins,
                         const float* dirs, int64_t n_rays, const int32_t* n_proc,
                         int64_t* counts, int32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif /* PXL_H */
