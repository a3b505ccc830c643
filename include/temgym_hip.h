/*
 * temgym_hip.h -- C ABI of libtemgym_hip.so, the MI355X (gfx950) kernels behind
 * temgymcore_amd's drop-in Ray / Component / run_to_end API.
 *
 * Conventions (every entry point):
 *   - Plain pointers to device memory, int64 sizes, POD parameter structs.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - Work is enqueued on `stream`; no entry point synchronises the device
 *     unless its comment says so.
 *   - The caller owns and allocates every input, output and workspace buffer.
 *   - Return value: 0 on success, a negative TG_E* code on failure.
 *     tg_last_error() returns a thread-local description of the last failure.
 *   - No C++ exception crosses this boundary.
 */
#ifndef TEMGYM_HIP_H
#define TEMGYM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TG_OK 0
#define TG_EINVAL (-1)   /* argument validation failed            */
#define TG_EHIP (-2)     /* a HIP runtime call failed              */
#define TG_ENOSYS (-3)   /* requested variant not implemented      */

/* Last error message of the calling thread ("" if none). */
const char* tg_last_error(void);

/* ABI version; bumped whenever a struct layout below changes. */
int32_t tg_abi_version(void);  /* currently 7 */

/* ======================================================================== *
 * Ray branch: per-ray register programs
 *
 * Replaces the element-wise op stream that run_iter / run_to_end issue for a
 * vectorised Ray (reference src/temgym_core/run.py:47-116, propagator.py:51-72,
 * components.py:133-523) and the numba binning kernel Grid.into_image ->
 * utils.inplace_sum (reference grid.py:231-280, utils.py:83-114).
 *
 * A program is a straight-line list of scalar instructions over a small
 * per-ray register file.  Instruction immediates hold host-folded constants
 * (parameters combined in double precision, then rounded once to the ray
 * dtype, as weakly typed Python scalars are in the reference).
 * ======================================================================== */

#define TG_MAX_INSTR 112
/* Programs whose registers all lie below 16 run on the kernel with a
 * 16-register file in VGPRs (4 f32 / 2 f64 rays per lane); wider ones (e.g.
 * the forward-mode tangents of nonlinear elements) on a kernel that keeps
 * TG_MAX_REG registers per ray in scratch memory (one ray per lane). */
#define TG_MAX_REG 64
#define TG_MAX_SLOT 32
#define TG_MAX_DIM 4

enum tg_opcode {
  TG_OP_ADD = 0,    /* d = a + b                          */
  TG_OP_SUB = 1,    /* d = a - b                          */
  TG_OP_MUL = 2,    /* d = a * b                          */
  TG_OP_DIV = 3,    /* d = a / b  (IEEE correctly rounded) */
  TG_OP_ADDI = 4,   /* d = a + imm                        */
  TG_OP_SUBI = 5,   /* d = a - imm                        */
  TG_OP_RSUBI = 6,  /* d = imm - a                        */
  TG_OP_MULI = 7,   /* d = a * imm                        */
  TG_OP_DIVI = 8,   /* d = a / imm                        */
  TG_OP_RDIVI = 9,  /* d = imm / a                        */
  TG_OP_NEG = 10,   /* d = -a                             */
  TG_OP_MOV = 11,   /* d = a                              */
  TG_OP_SET = 12,   /* d = imm                            */
  TG_OP_SIGN = 13,  /* d = sign(a): -1, 0, +1, NaN->NaN    */
  TG_OP_COUNT = 14,
  /* Flag or-ed into `op` in TG_F64 programs: the operation's operands and
   * result are f32 values (JAX-style per-operation promotion of mixed f32/f64
   * bundles).  The f64 kernel computes in f64 and rounds the result to f32,
   * which is the correctly rounded f32 result for + - * / (53 >= 2*24+2). */
  TG_OP_F32 = 0x80
};

enum tg_dtype { TG_F32 = 0, TG_F64 = 1 };

typedef struct tg_instr {
  uint8_t op, dst, a, b;
  float immf;   /* imm rounded to f32 (used by TG_F32 programs) */
  double imm;   /* imm in f64 (used by TG_F64 programs)         */
} tg_instr;     /* 16 bytes */

/* One varying operand: a device array broadcast over the logical shape.
 * element offset of multi-index (i0..i3) = sum_k idx_k * stride[k]. */
typedef struct tg_slot {
  const void* ptr;
  int64_t stride[TG_MAX_DIM];
  int32_t reg;      /* register the value is loaded into */
  int16_t contig;   /* 1: element i of the flattened shape is ptr[i] */
  int16_t dtype;    /* tg_dtype of the array (f32 arrays may feed f64 programs) */
} tg_slot;

typedef struct tg_program {
  int32_t dtype;             /* tg_dtype                                   */
  int32_t n_instr;
  int32_t n_slot;
  int32_t n_out;
  int64_t n;                 /* number of rays = prod(shape)               */
  int64_t shape[TG_MAX_DIM]; /* logical broadcast shape, leading dims = 1  */
  tg_instr instr[TG_MAX_INSTR];
  tg_slot slot[TG_MAX_SLOT];
  void* out[TG_MAX_SLOT];    /* contiguous outputs of n elements           */
  int32_t out_reg[TG_MAX_SLOT];
  int32_t out_dtype[TG_MAX_SLOT]; /* tg_dtype of each output array         */
} tg_program;

/* Detector binning fused after a program (or run alone with n_instr = 0).
 * pixel (py, px) = round_half_even(M @ [y, x, 1]) computed in f64, then the
 * reference's int32 cast (NaN -> 0, saturating) and 0 <= p < shape bounds. */
typedef struct tg_bin_spec {
  int32_t enabled;
  int32_t acc_dtype;         /* 0 = int32, 1 = int64, 2 = float32, 3 = float64 */
  int32_t x_reg, y_reg;      /* registers holding x and y (or -1: uniform)  */
  double x_uniform, y_uniform;
  double m[6];               /* metres->pixels rows: py = m0*y + m1*x + m2; px = m3*y + m4*x + m5 */
  int64_t height, width;
  void* acc;                 /* [height, width] accumulator, row-major      */
  int32_t w_reg;             /* weight register (-1: weight 1)              */
  int32_t grid_hint;         /* workgroups to launch (0: built-in heuristic) */
} tg_bin_spec;

/* Run `prog` over prog->n rays; optionally bin the final positions.
 * Ray-count 0 is a no-op. */
int tg_ray_program_run(const tg_program* prog, const tg_bin_spec* bin, void* stream);

/* Number of compute units of the current device (cached per device). */
int32_t tg_device_cu_count(void);

/* ---- autotuning ----------------------------------------------------------
 * A per-process table of launch choices, keyed by (knob, current device,
 * key[4]).  The autotuner (temgymcore_amd/autotune.py) times candidates and
 * records the fastest; the entry points below consult the table and fall back
 * to built-in heuristics.  Thread-safe.  No reference counterpart (the
 * reference leaves launch shapes to XLA).
 *   TG_TUNE_BEAMLET_SPLITS key {n, H, W, 0}: beamlet slices of tg_beamlet_sum
 *                          (and of tg_beamlet_sum_workspace), 1..128
 *   TG_TUNE_RAY_GRID       key {n, H*W, dtype, 0}: workgroups of a binning
 *                          tg_ray_program_run whose bin->grid_hint is 0
 *   TG_TUNE_FIT_SLICE      key {n_inst, n_gauss, n_pix, 0}: target Gaussians
 *                          per forward slice in tg_gauss_fit_plan, 16..4096
 *   TG_TUNE_STEM_FIT       key {S, J, n_cam, 0}: tg_stem_fit launch variant
 *                          (tg_stem_fit_variant)
 *   TG_TUNE_BEAMLET_PATH   key {n, H, W, 0}: tg_beamlet_sum's tile sum
 *                          (tg_beamlet_path)
 *   TG_TUNE_TILE_SHAPE     tile shape of the VALU recurrence tile sum
 *                          (tg_tile_shape): key {n, H, W, 0} for
 *                          tg_beamlet_sum, {n_inst, n_gauss, n_pix, 1} for
 *                          tg_gauss_fit_plan (the plan records it) */
enum tg_tune_knob {
  TG_TUNE_BEAMLET_SPLITS = 1,
  TG_TUNE_RAY_GRID = 2,
  TG_TUNE_FIT_SLICE = 3,
  TG_TUNE_STEM_FIT = 4,
  TG_TUNE_BEAMLET_PATH = 5,
  TG_TUNE_TILE_SHAPE = 6
};
/* Tile of the VALU recurrence tile sum (256 threads; each lane walks a run of
 * pixels of one row): width x height, run length.  Default 64x64, run 16. */
enum tg_tile_shape {
  TG_TILE_64x64_RUN16 = 1,
  TG_TILE_128x32_RUN16 = 2,
  TG_TILE_32x128_RUN16 = 3,
  TG_TILE_64x32_RUN8 = 4
};
/* TG_BEAMLET_PATH_AUTO: the matrix-core tile sum when every beamlet of the
 * launch is separable (no px*py term in Phi, envelope decaying along both
 * axes), else the VALU recurrence; TG_BEAMLET_PATH_VALU: always the VALU
 * recurrence. */
enum tg_beamlet_path { TG_BEAMLET_PATH_AUTO = 1, TG_BEAMLET_PATH_VALU = 2 };
int tg_tune_set(int32_t knob, const int64_t* key, int32_t value);
/* 1 and *value set when an entry exists, 0 when not, < 0 on error. */
int tg_tune_get(int32_t knob, const int64_t* key, int32_t* value);
int tg_tune_clear(void);

/* ======================================================================== *
 * 4D-STEM shadow-image model (scan x camera-length batches)
 *
 * Per scan position s and ray j the model is the component chain
 *   point source at z=0 (x = y = 0, slopes on an n_side x n_side grid)
 *   -> Scanner(z=0, sp, st)  -> sample plane at z=defocus
 *   -> descanner with a 4x4 error map of (spx, spy, stx, sty) plus offsets
 *   -> Detector at z = defocus + L_c,
 * with scan positions from a rotated ScanGrid (reference components.py:218-372,
 * coordinate_transforms.py:50-101) and scan tilt st = pivot * sp.
 * The descanner reduces to the reference Descanner (components.py:320-372)
 * when the tilt columns of the 4x4 map are zero.  Detector coordinates are
 * affine in the ray slopes for every (s, c), which the kernels exploit.
 *
 * Parameter vector (device, f64, 18 + n_cam entries):
 *   [0..15]  descan map A, row-major: rows (x, y, dx, dy) of the output
 *            offset, columns (spx, spy, stx, sty)
 *   [16]     scan step (m)      [17] scan rotation (degrees)
 *   [18+c]   camera length L_c (m)
 * ======================================================================== */

#define TG_STEM_MAX_CAM 16

typedef struct tg_stem_geometry {
  int32_t scan_h, scan_w;   /* scan grid shape (rows, cols)                    */
  int32_t n_side;           /* rays per side; J = n_side^2 (n_side even)      */
  int32_t n_cam;            /* camera lengths (1..TG_STEM_MAX_CAM)            */
  double semi_conv;         /* slopes (k - (n_side-1)/2) * 2*semi_conv/n_side  */
  double defocus;           /* source -> sample plane distance (m)            */
  double pivot;             /* scan tilt per metre of scan position (rad/m)    */
  double offsets[4];        /* descan offsets added to x, y, dx, dy           */
  int32_t img_h, img_w;     /* sample image shape                              */
  double img_m[6];          /* metres -> fractional sample pixel (py, px)      */
  int32_t det_h, det_w;     /* detector shape                                  */
  double det_m[6];          /* metres -> detector pixel (py, px)               */
} tg_stem_geometry;

/* Forward model for every (camera length, scan position, ray).
 *   coords  [n_cam, S, J, 2] f32 detector (x, y) or NULL
 *   values  [S, J] f32 bilinear sample-image value at the sample plane or NULL
 *   image4d [S, det_h, det_w] f32 (accumulated, camera length 0) or NULL:
 *           each ray adds its sample value at its detector pixel (binning
 *           semantics of Grid.into_image).  The caller zeroes it.
 * `image` is the [img_h, img_w] f32 sample; required if values or image4d. */
int tg_stem_forward(const tg_stem_geometry* g, const double* params, const float* image, float* coords,
                    float* values, float* image4d, void* stream);

/* Descan / geometry fit, forward and adjoint in one pass over the targets.
 *   targets  [n_cam, S, J, 2] f32 target detector coordinates
 *   workspace tg_stem_fit_workspace(g) bytes of f64: starts with the
 *            [S, 18 + n_cam + 1] per-scan-position terms (one row per
 *            scan position, one column per output entry), then
 *            [S, n_cam, 7] per-camera moments
 *   out      [18 + n_cam + 1] f64: d loss / d params, then the loss
 * loss = sum over (c, s, j) of |detector(x, y) - target|^2.
 * Two launch variants give identical results: TG_STEM_FIT_SCAN (one
 * workgroup per scan position; the default when S >= 4 x the CU count) and
 * TG_STEM_FIT_CAMERA (one per scan position and camera length; the default
 * for smaller scans); the tuning knob TG_TUNE_STEM_FIT overrides. */
/* Shadow-image back-projection: the adjoint of tg_stem_forward's 4D deposit
 * with respect to the sample image (the consumer of 4D-STEM data the
 * reference describes, README.md:13-16).  For every ray (s, j) whose detector
 * pixel for camera `cam` (binned as the deposit bins it) lies on the
 * detector, image[q] += data4d[s, pixel] * w_q and, if weight is not NULL,
 * weight[q] += w_q, where w_q are the bilinear weights of sample pixel q at
 * the ray's sample-plane position.  data4d [S, det_h, det_w] f32; image and
 * weight [img_h, img_w] f64, accumulated into (the caller zeroes them).  The
 * FP64 sums are order-dependent only at the last ulp (atomics).  One scan
 * position's detector holds at most 2^30 pixels (TG_EINVAL beyond). */
int tg_stem_backproject(const tg_stem_geometry* g, const double* params, const float* data4d, int32_t cam,
                        double* image, double* weight, void* stream);

/* The launches of tg_stem_fit before its final reduction: fills the head of
 * the workspace with the [S, 18 + n_cam + 1] per-scan-position terms (row s,
 * column t: scan position s's contribution to out[t]; the last column is its
 * share of the loss, a residual map over the scan).  tg_stem_fit = this +
 * the fixed-order FP64 sums over s. */
int tg_stem_fit_terms(const tg_stem_geometry* g, const double* params, const float* targets, double* workspace,
                      void* stream);

enum tg_stem_fit_variant { TG_STEM_FIT_SCAN = 1, TG_STEM_FIT_CAMERA = 2 };
int64_t tg_stem_fit_workspace(const tg_stem_geometry* g);
int tg_stem_fit(const tg_stem_geometry* g, const double* params, const float* targets, double* workspace,
                double* out, void* stream);

/* tg_stem_fit without a workspace argument and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, allocation-free serving and
 * fitting loops).  The workspace is a device array inside the library, one
 * per GPU, held until the process ends, of tg_stem_fit_noalloc_max_bytes()
 * bytes (8 MiB: every 64x64 scan up to TG_STEM_MAX_CAM camera lengths); a
 * geometry whose tg_stem_fit_workspace(g) exceeds it returns TG_EINVAL before
 * any launch.  Same kernels and results as tg_stem_fit.  `stream` must belong
 * to the current device.  Calls on one GPU share its array: calls that overlap
 * on one GPU (other streams, other threads) race and must be serialised by
 * the caller. */
int64_t tg_stem_fit_noalloc_max_bytes(void);
int tg_stem_fit_noalloc(const tg_stem_geometry* g, const double* params, const float* targets, double* out,
                        void* stream);

/* ======================================================================== *
 * Gaussian beamlets (reference gaussian.py:225-427)
 *
 * Every beamlet contributes exp(Phi(px, py)) to detector pixel (py, px),
 * where Phi is a complex quadratic in the pixel indices obtained by folding
 * the reference's per-beamlet 2x2 complex algebra (_beam_field,
 * gaussian.py:276-316, or _input_beam_field, gaussian.py:402-407), the
 * amplitude prefactor, the k*pathlength phase and the pixel->metre transform
 * of the Grid into six complex coefficients (computed in f64).
 * ======================================================================== */

/* Per-beamlet parameter row (f64, TG_BEAMLET_NPARAM columns):
 * x, y, dx, dy, pathlength, amp_re, amp_im, waist_x, waist_y, radius_x,
 * radius_y (inf allowed), wavelength, theta, 0, 0, 0. */
#define TG_BEAMLET_NPARAM 16

enum tg_beamlet_mode {
  TG_BEAMLET_PROPAGATED = 0,  /* make_gaussian_image: field after a 5x5 ABCD   */
  TG_BEAMLET_INPUT = 1        /* evaluate_gaussian_input_image: input plane    */
};

typedef struct tg_beamlet_coef {
  double phi[12];      /* (re, im) of a, b, c, d, e, f:
                          Phi = a px^2 + b px py + c py^2 + d px + e py + f      */
  float k_re, k_im;    /* exp(2a): ratio of successive x-step ratios             */
  float cx, cy;        /* envelope centre (pixels)                               */
  /* kept pixels of row py (v = py - cy): cx - alpha v +- sqrt(beta - gamma v^2)
   * (gamma <= 0: the whole row); the cutoff ellipse is
   * (p-c)^T S (p-c) <= cut with |field| = peak * exp(-(p-c)^T S (p-c)) */
  float alpha, beta, gamma;
  float cut;
  float box_w, box_h;  /* half-extents of that ellipse (pixels; huge: unbounded) */
} tg_beamlet_coef;     /* 136 bytes */

/* Coefficients of n beamlets.
 *   params  [n, TG_BEAMLET_NPARAM] f64 (device)
 *   abcd    [n or 1, 5, 5] f64 ray-transfer matrices (propagated mode; NULL
 *           in input mode); abcd_stride = 25 (per beamlet) or 0 (shared)
 *   pix2m   [6] f64: y_m = m0*py + m1*px + m2, x_m = m3*py + m4*px + m5
 *   nsigma  cutoff: contributions below exp(-nsigma^2) of the beamlet's peak
 *           amplitude are dropped (4 = the "4 sigma" cut; <= 0: keep down to
 *           exp(-80), i.e. below f32 resolution). */
int tg_beamlet_coefficients(const double* params, int64_t n, const double* abcd, int64_t abcd_stride, int32_t mode,
                            const double* pix2m, double nsigma, tg_beamlet_coef* out, void* stream);

/* field[H, W] (complex64 as interleaved f32 pairs) = sum over beamlets of
 * exp(Phi_b(px, py)), overwritten.  `workspace` needs
 * tg_beamlet_sum_workspace(n, H, W) bytes.
 * Two tile sums (tg_beamlet_path), chosen on the device per launch:
 *   - every beamlet separable (Phi without a px*py term, envelope decaying
 *     along both axes, as for untilted round or axis-aligned beams): each
 *     64x64 tile is a complex matrix product Ey^T Ex over the beamlets whose
 *     cutoff box overlaps it, Ex/Ey from f64 phases, on the matrix cores
 *     (f32 MFMA); culling is per tile, so terms below the cut inside an
 *     overlapping tile are summed too;
 *   - otherwise the per-lane f32 recurrence over 16-pixel runs, culled per
 *     run to the cutoff ellipse. */
int64_t tg_beamlet_sum_workspace(int64_t n, int64_t height, int64_t width);
int tg_beamlet_sum(const tg_beamlet_coef* coef, int64_t n, int64_t height, int64_t width, float* field,
                   void* workspace, int64_t workspace_bytes, void* stream);

/* Gradient of a real loss L of that field with respect to the beamlet
 * parameters (and, in propagated mode, the ray-transfer matrices): the
 * backward of tg_beamlet_coefficients + tg_beamlet_sum, for make_gaussian_image
 * / evaluate_gaussian_input_image under jax.grad (gaussian.py:225-427).
 *   params, n, abcd, abcd_stride, mode, pix2m, nsigma: as for the forward
 *   grad_field  [H, W] complex64: dL/dRe(field) + i dL/dIm(field)
 *   grad_params [n, TG_BEAMLET_NPARAM] f64 (columns 13-15 are 0; the
 *               amplitude as its real and imaginary parts)
 *   grad_abcd   [n, 25] f64 per-beamlet dL/dT (row-major; the constant last
 *               row is 0), propagated mode only, or NULL.  With a shared
 *               matrix (abcd_stride 0) its gradient is the sum over beamlets.
 * `workspace` needs tg_beamlet_backward_workspace(n, H, W) bytes, 256-byte
 * aligned. */
int64_t tg_beamlet_backward_workspace(int64_t n, int64_t height, int64_t width);
int tg_beamlet_backward(const double* params, int64_t n, const double* abcd, int64_t abcd_stride, int32_t mode,
                        const double* pix2m, double nsigma, int64_t height, int64_t width, const float* grad_field,
                        double* grad_params, double* grad_abcd, void* workspace, int64_t workspace_bytes,
                        void* stream);

/* ======================================================================== *
 * Batched Gaussian-decomposition fits
 *
 * Many independent instances, each a set of input-plane Gaussians
 * (evaluate_gaussian_input_image, gaussian.py:372-427) sampled on its own
 * detector grid, with the intensity loss of the reference's fitting example
 * (examples/fit_three_gaussians.ipynb, cell 2)
 *     L_i = mean over pixels of (|u_i|^2 - T_i)^2
 * and its gradient with respect to every parameter row (the jax.grad of
 * that loss in the example).  Instance sizes may differ; all instances run
 * in the same launches.
 * ======================================================================== */

typedef struct tg_fit_instance {
  int64_t gauss_begin;  /* first parameter row of the instance                 */
  int64_t n_gauss;      /* its number of Gaussians (0 allowed)                 */
  int64_t pix_begin;    /* first pixel of its image in the packed pixel arrays */
  int32_t height, width;
  double pix2m[6];      /* pixel->metre map, as for tg_beamlet_coefficients    */
} tg_fit_instance;      /* 80 bytes */

/* Launch geometry and workspace layout, filled in by tg_gauss_fit_plan. */
typedef struct tg_fit_plan {
  int32_t n_inst, n_items, n_loss_blocks, n_moment_blocks;
  int64_t n_gauss;          /* parameter rows                                     */
  int64_t n_assigned;       /* rows that belong to an instance                    */
  int64_t n_pix;            /* length of the packed pixel arrays                  */
  float cut;
  int32_t tile_shape;       /* tg_tile_shape of the forward tile sum              */
  int64_t off_inst, off_items, off_loss_blocks, off_order, off_gauss_inst;
  int64_t off_coef, off_unit, off_partial, off_weight, off_loss_part, off_moments;
  int64_t table_bytes;      /* bytes of the tables uploaded by tg_gauss_fit_upload */
  int64_t workspace_bytes;  /* total device workspace                             */
} tg_fit_plan;

/* Host-only: checks the instances (Gaussian rows inside [0, n_gauss), pixel
 * and Gaussian ranges pairwise disjoint) and lays out the workspace.  nsigma
 * as for tg_beamlet_coefficients.  No reference counterpart (the reference
 * fits one instance at a time). */
int tg_gauss_fit_plan(const tg_fit_instance* inst, int32_t n_inst, int64_t n_gauss, double nsigma,
                      tg_fit_plan* plan);

/* Copies the plan's tables (built from the same host `inst` array) into the
 * device workspace of plan->workspace_bytes and waits for the copy. */
int tg_gauss_fit_upload(const tg_fit_plan* plan, const tg_fit_instance* inst, void* workspace, void* stream);

/* One pass of every instance, stream-ordered, no allocation, no host sync:
 *   params  [n_gauss, TG_BEAMLET_NPARAM] f64 parameter rows
 *   target  [n_pix] f32 target intensities |u_target|^2, or NULL
 *   field   [n_pix] complex64 (interleaved f32) fields, or NULL
 *   loss    [n_inst] f64 losses (requires target), or NULL
 *   grad    [n_gauss, TG_BEAMLET_NPARAM] f64 d(sum of losses)/d(params)
 *           (requires target; columns 13-15 and rows of no instance are 0),
 *           or NULL
 * Replaces evaluate_gaussian_input_image plus the jax.grad of the example's
 * loss (gaussian.py:372-427; fit_three_gaussians.ipynb cells 2-3). */
int tg_gauss_fit_run(const tg_fit_plan* plan, void* workspace, const double* params, const float* target,
                     float* field, double* loss, double* grad, void* stream);

/* ======================================================================== *
 * Measurement support
 * ======================================================================== */

/* Streams `bytes` (multiple of 16) from `src` with the loads the streaming
 * kernels use and writes one partial sum per workgroup to block_sums[n_blocks].
 * Used by bench.py to report the achievable HBM read rate of the device.
 * No reference counterpart (measurement only). */
int tg_hbm_read_probe(const void* src, int64_t bytes, float* block_sums, int32_t n_blocks, void* stream);

/* ======================================================================== *
 * Ray sources generated on the device (reference source.py:82-188,
 * utils.py:46-80, :117-205, :297-325).  Outputs are n-element device arrays
 * of `dtype` (tg_dtype); tables are device arrays.
 * kind 0 = PointSource: x = off_x, y = off_y, (dx, dy) = (r cos a, r sin a);
 * kind 1 = ParallelBeam: (x, y) = (r cos a + off_x, r sin a + off_y), dx = dy = 0.
 * ======================================================================== */

/* Deterministic concentric rings: replaces concentric_rings with its
 * multi_cumsum_inplace angles (utils.py:117-175, :46-80) behind
 * generate_array(random=False).  ring_radius/ring_step [n_rings]: radius and
 * angle step of each ring; ring_first [n_rings]: first point of each ring
 * (ring_first[0] = 0); seg_first [n_seg]: first point of each run of the
 * reference's cumulative angle sum (seg_first[0] = 0).  Point v gets the f64
 * sum, in order, of the steps of points seg+1..v of its run. */
int tg_source_rings(const double* ring_radius, const double* ring_step, const int64_t* ring_first, int32_t n_rings,
                    const int64_t* seg_first, int32_t n_seg, int64_t n, int32_t kind, double off_x, double off_y,
                    int32_t dtype, void* x, void* y, void* dx, void* dy, void* stream);

/* Fibonacci spiral (utils.py:297-325): r_i = radius sqrt((i + 0.5) / denom)
 * for i <= band, radius beyond, r_0 = 0; phi_i = i * ga;
 * x = r cos phi, y = r sin phi. */
int tg_source_fibonacci(int64_t n, double radius, double band, double denom, double ga, int32_t dtype, void* x,
                        void* y, void* stream);

/* Uniform points in a disc of radius `scale` (replaces random_coords,
 * utils.py:178-205, behind generate_array(random=True)): exactly n points from
 * a counter-based Philox4x32-10 stream keyed by `seed`, r = scale sqrt(u). */
int tg_source_disc(int64_t n, uint64_t seed, double scale, int32_t kind, double off_x, double off_y, int32_t dtype,
                   void* x, void* y, void* dx, void* dy, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TEMGYM_HIP_H */
