/*
 * dvr.h -- C-ABI of the MI355X differentiable volume renderer (libdvr_hip.so).
 *
 * Plain pointers, sizes and strides only; no framework types cross this
 * boundary.  Every entry point returns a dvr_status; on failure
 * dvr_last_error() returns a thread-local, human-readable message (the
 * reference raises c10::Error with the same text through pybind11).
 *
 * Semantics follow DiffDVR's renderer (paths relative to that repository):
 *   - tensors may have arbitrary element strides; a batch dimension of size 1
 *     broadcasts over the render batch (renderer/renderer_tensor.cuh:122-269,
 *     renderer/pytorch_utils.h:12-20);
 *   - the caller allocates every output; adjoint outputs are accumulated into
 *     and must be zero-initialised by the caller
 *     (renderer/renderer_settings.cuh:227-228);
 *   - launches are asynchronous on the given HIP stream (NULL = default stream).
 *
 * All floating-point tensors are float32 (the reference's real_t with
 * USE_DOUBLE_PRECISION=0), except that the elementwise losses also take
 * float64 (AT_DISPATCH_FLOATING_TYPES in the reference); terminationIndex is
 * int32.
 */
#ifndef DVR_H_
#define DVR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dvr_status;
#define DVR_OK 0
#define DVR_ERR_INVALID_ARGUMENT 1 /* validation failure (reference: TORCH_CHECK) */
#define DVR_ERR_NO_GRADIENT 2      /* no adjoint output enabled (renderer.cpp:712-717) */
#define DVR_ERR_UNSUPPORTED 3      /* combination not available on this device path */
#define DVR_ERR_HIP 4              /* HIP runtime / launch error */

/* dtypes */
#define DVR_FLOAT32 0
#define DVR_INT32 1
#define DVR_FLOAT64 2 /* the elementwise losses only */
#define DVR_BOOL 3    /* one byte per element, 0 or 1 (importance-sampling output) */

/* Enum values are those of renderer/renderer_settings.cuh:81-146 and of the
 * pyrenderer enums (bindings/bindings.cpp:215-232). */
#define DVR_FILTER_NEAREST 0
#define DVR_FILTER_TRILINEAR 1
#define DVR_FILTER_PRESHADED 3 /* volume = 4 x X x Y x Z colours (r, g, b, absorption); needs DVR_TF_PRESHADED */
#define DVR_CAMERA_RAY_START_DIR 0
#define DVR_CAMERA_INVERSE_VIEW_MATRIX 1
#define DVR_CAMERA_REFERENCE_FRAME 2
#define DVR_TF_IDENTITY 0
#define DVR_TF_TEXTURE 1
#define DVR_TF_LINEAR 2
#define DVR_TF_GAUSSIAN 3
#define DVR_TF_PRESHADED 5     /* the TF passes the pre-shaded colour on; any TF tensor size */
#define DVR_TF_GAUSSIAN_LOG 6  /* log of the Gaussian TF: pre-shading only, not a render mode (as the reference) */
/* R*C entries of the largest TF the tiled, checkpointed, fused, forward-mode and batched kernels
 * stage in LDS; larger TFs render on the plain path (forward and inversion adjoint) only. */
#define DVR_MAX_TILED_TF_ENTRIES 4096
#define DVR_BLEND_BEER_LAMBERT 0
#define DVR_BLEND_ALPHA 1

/* A strided view of device (or, where stated, host) memory. */
typedef struct dvr_tensor {
    void* data;         /* NULL with ndim == 0 = not given (an empty tensor has ndim >= 1) */
    int64_t sizes[6];
    int64_t strides[6]; /* in elements */
    int32_t ndim;
    int32_t dtype;      /* DVR_FLOAT32 / DVR_INT32 / DVR_FLOAT64 / DVR_BOOL */
    int32_t device;     /* HIP device ordinal, -1 = host memory */
} dvr_tensor;

/* Mirrors renderer::RendererInputsHost (renderer/renderer.h:20-64).
 * box_min, box_size, step_size and an InverseViewMatrix camera can be given
 * either as a tensor or as a value (the std::variant alternatives of the
 * reference): the *_value field is used when the tensor's data is NULL. */
typedef struct dvr_inputs {
    int32_t screen_width, screen_height;
    int32_t volume_filter_mode;
    dvr_tensor volume;                       /* (B) x X x Y x Z */
    dvr_tensor box_min;  float box_min_value[3];   /* (B) x 3 */
    dvr_tensor box_size; float box_size_value[3];  /* (B) x 3 */
    int32_t camera_mode;
    dvr_tensor camera_ray_start;             /* (B) x H x W x 3   (RayStartDir) */
    dvr_tensor camera_ray_dir;               /* (B) x H x W x 3   (RayStartDir) */
    dvr_tensor camera_matrix;                /* (B) x 4 x 4 (InverseViewMatrix) or (B) x 3 x 3 (ReferenceFrame) */
    float camera_matrix_value[16];           /* row-major 4x4, InverseViewMatrix given as a value (glm::mat4) */
    float camera_fov_y_radians;              /* ReferenceFrame only */
    dvr_tensor step_size; float step_size_value;   /* (B) x (H) x (W) */
    int32_t tf_mode;
    dvr_tensor tf;                           /* (B) x R x C */
    int32_t blend_mode;
    float blending_early_out;                /* default 1 - 1e-5 */
} dvr_inputs;

/* Mirrors renderer::RendererOutputsHost (renderer/renderer.h:67-71). */
typedef struct dvr_outputs {
    dvr_tensor color;             /* B x H x W x 4 */
    dvr_tensor termination_index; /* B x H x W, int32 */
} dvr_outputs;

/* Mirrors renderer::AdjointOutputsHost (renderer/renderer.h:95-111). */
typedef struct dvr_adjoint_outputs {
    int32_t has_volume_derivatives;   dvr_tensor adj_volume;          /* (B) x X x Y x Z */
    int32_t has_stepsize_derivatives; dvr_tensor adj_stepsize;        /* (B) x (H) x (W) */
    int32_t has_camera_derivatives;   dvr_tensor adj_camera_ray_start;/* (B) x H x W x 3 */
                                      dvr_tensor adj_camera_ray_dir;  /* (B) x H x W x 3 */
    int32_t has_tf_derivatives;       dvr_tensor adj_tf;              /* (B) x R x C */
    int32_t tf_delayed_accumulation;
} dvr_adjoint_outputs;

/* Kernel family selection. */
#define DVR_PATH_AUTO 0        /* tiled path where available, else plain */
#define DVR_PATH_PLAIN 1       /* reference structure: one thread per pixel, row order, grid-stride */
#define DVR_PATH_TILED 2       /* 8x8 ray tiles per wave64, XCD-aware tile order */

#define DVR_ADJOINT_AUTO 0       /* checkpointed */
#define DVR_ADJOINT_INVERSION 1  /* reference: reconstruct by inverting the blend step */
#define DVR_ADJOINT_CHECKPOINT 2 /* store alpha every K samples, recompute segments */

typedef struct dvr_forward_options {
    int32_t path;                /* DVR_PATH_* */
    int32_t checkpoint_interval; /* K; 0 = no checkpoints written */
    dvr_tensor checkpoints;      /* float32 buffer of >= dvr_checkpoint_buffer_elements() */
    dvr_tensor densities;        /* optional, with checkpoints: float32 buffer of >= dvr_density_buffer_elements()
                                    that receives every sample's density for the checkpointed adjoint */
} dvr_forward_options;

typedef struct dvr_adjoint_options {
    int32_t path;                /* DVR_PATH_* */
    int32_t method;              /* DVR_ADJOINT_* */
    int32_t checkpoint_interval; /* K in {8,16,32,64}; 0 = default (16) */
    dvr_tensor checkpoints;      /* optional: written by dvr_render_forward_ex with the same K and image batch
                                    (checked: the library keeps a host-side stamp of every checkpoint buffer a
                                    forward writes -- K, B x H x W and the density buffer it filled -- and
                                    rejects a mismatch with DVR_ERR_INVALID_ARGUMENT) */
    dvr_tensor densities;        /* optional, with checkpoints: the densities that same forward call wrote; the
                                    reverse march reads them instead of re-interpolating the volume (without
                                    checkpoints the library uses a scratch store of its own when it fits 2 GiB) */
} dvr_adjoint_options;

/* ---- library ---------------------------------------------------------- */

/*
 * Library scratch: the checkpointed adjoint without caller checkpoints, fused
 * rendering, and the TF tables of the forward-mode entry points use device
 * buffers the library keeps per device (grown on demand: a call that needs a
 * larger one synchronises the device, frees the old buffer and allocates).
 * Calls on different streams of one device are ordered: a call on another
 * stream than the previous user of the scratch makes its stream wait for an
 * event recorded after that use, and calls from several host threads take
 * turns until their work is enqueued.  Under HIP graph capture this is not
 * tracked: a captured graph's uses of the scratch are ordered by the graph and
 * by the stream it is replayed on only, so its replays must not overlap other
 * calls that use the scratch.  Entry points given every buffer by the caller
 * (checkpoints, densities) and batch plans, which own their scratch, have no
 * such restriction.
 */

/* Thread-local message of the last failed call ("" if none). */
const char* dvr_last_error(void);
/* 0: this library computes in float32 (bindings/bindings.cpp:143-146). */
int32_t dvr_use_double_precision(void);
/* hipDeviceSynchronize (bindings/bindings.cpp:306). */
dvr_status dvr_sync(void);
/* Version string of the library build. */
const char* dvr_version(void);

/* ---- rendering (replaces renderer::Renderer, renderer/renderer.h:159-268) -- */

/* Renderer::renderForward (renderer/renderer.h:159-161; renderer/renderer.cpp:393-417). */
dvr_status dvr_render_forward(const dvr_inputs* inputs, dvr_outputs* outputs, void* stream);

/* renderForward with a kernel-family choice and optional alpha checkpoints
 * for a later checkpointed adjoint.  No reference counterpart. */
dvr_status dvr_render_forward_ex(const dvr_inputs* inputs, dvr_outputs* outputs,
                                 const dvr_forward_options* options, void* stream);

/* Renderer::renderAdjoint (renderer/renderer.h:264-268; renderer/renderer.cpp:660-742). */
dvr_status dvr_render_adjoint(const dvr_inputs* inputs, const dvr_outputs* outputs_from_forward,
                              const dvr_tensor* adj_color, dvr_adjoint_outputs* adj_outputs, void* stream);

/* renderAdjoint with method / path / checkpoint choices. */
dvr_status dvr_render_adjoint_ex(const dvr_inputs* inputs, const dvr_outputs* outputs_from_forward,
                                 const dvr_tensor* adj_color, dvr_adjoint_outputs* adj_outputs,
                                 const dvr_adjoint_options* options, void* stream);

/* Per-pixel image loss whose gradient the fused kernels form in registers. */
#define DVR_LOSS_L1 1 /* mean |color - target|   (torch.nn.L1Loss)  */
#define DVR_LOSS_L2 2 /* mean (color - target)^2 (torch.nn.MSELoss) */
typedef struct dvr_loss {
    int32_t kind;       /* DVR_LOSS_L1 / DVR_LOSS_L2 */
    dvr_tensor target;  /* B x H x W x 4 */
    float scale;        /* factor of the summed loss; 0 = 1/(B*H*W*4), i.e. the mean */
    dvr_tensor loss;    /* optional 1-element float32: the loss value is ADDED to it */
} dvr_loss;

/* Forward march, image loss against `target` and checkpointed adjoint in ONE
 * launch per batch: colour and termination index are written to `outputs`,
 * the loss gradient never leaves registers, and the adjoint outputs are
 * accumulated as by dvr_render_adjoint.  checkpoint_interval K in {0,8,16,32,64}
 * (0 = default).  No reference counterpart: it fuses Renderer::renderForward,
 * the caller's torch loss (pytests/tests/volume/compare_reconstruction.py:359-366)
 * and Renderer::renderAdjoint. */
dvr_status dvr_render_fused(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                            dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval, void* stream);

/* dvr_render_fused without any allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as dvr_render_fused (the reverse march
 * re-interpolates instead of reading a density store: the same densities), with the checkpoints in
 * a static per-device array of DVR_NOALLOC_SCRATCH_FLOATS floats (32 MiB, held by every GPU the
 * library runs on until the process ends).  Inputs whose checkpoints need more
 * (dvr_checkpoint_buffer_elements) return DVR_ERR_INVALID_ARGUMENT: use a larger
 * checkpoint_interval or fewer pixels per call.  The first call on a device sets the array up
 * (hipGetSymbolAddress, which makes the runtime load it): make that call before capturing a graph.
 * Calls on other streams of the GPU are ordered as for the library scratch above; replays of a
 * captured graph are not, and race with calls that overlap them on that GPU. */
#define DVR_NOALLOC_SCRATCH_FLOATS (8 * 1024 * 1024)
dvr_status dvr_render_fused_noalloc(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                                    dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval, void* stream);

/* Number of float32 elements a checkpoint buffer needs for these inputs
 * and interval K (0 if K == 0). */
int64_t dvr_checkpoint_buffer_elements(const dvr_inputs* inputs, int32_t checkpoint_interval);
/* Elements of the density buffer of dvr_render_forward_ex / dvr_render_adjoint_ex (one float per
 * sample of the longest ray, per pixel of every 8x8 tile); -1 on invalid inputs.  No reference
 * counterpart. */
int64_t dvr_density_buffer_elements(const dvr_inputs* inputs);

/* ---- forward-mode differentiation ---------------------------------------- */

#define DVR_NOT_ENABLED (-1)

/* Mirrors renderer::ForwardDifferencesSettingsHost (renderer/renderer.h:74-93):
 * which derivative index 0 <= d < D traces which scalar input. */
typedef struct dvr_forward_differences {
    int32_t D;                  /* number of traced derivatives, > 0 */
    int32_t d_stepsize;         /* index of the step size or DVR_NOT_ENABLED */
    int32_t d_ray_start[3];     /* indices of the evaluated ray start (x,y,z) */
    int32_t d_ray_dir[3];       /* indices of the evaluated, normalised ray direction */
    int32_t has_tf_derivatives;
    dvr_tensor d_tf;            /* int32, same shape as the TF: index of each entry or -1 */
} dvr_forward_differences;

/* Renderer::renderForwardGradients (renderer/renderer.h:189-196;
 * renderer/renderer.cpp:419-498).  Writes the colour and termination index
 * (equal to dvr_render_forward) and gradients_out (B x H x W x D x 4), the
 * derivative of every pixel's RGBA w.r.t. every traced input.  Indices must
 * lie in [0, D) and no two TF entries may share one index.  Volume
 * derivatives are not available in forward mode (as in the reference). */
dvr_status dvr_render_forward_gradients(const dvr_inputs* inputs, const dvr_forward_differences* settings,
                                        dvr_outputs* outputs, dvr_tensor* gradients_out, void* stream);

/* Renderer::forwardVariablesToGradients (renderer/renderer.h:198-215;
 * renderer/renderer.cpp:500-560): contracts forward_variables (B x H x W x D x 4)
 * with adj_color (B x H x W x 4) into the adjoint outputs for step size,
 * camera rays and TF.  All outputs are accumulated into (the reference
 * overwrites the step and camera entries; with the zero-initialised outputs
 * it is called with the results agree); outputs that broadcast over a
 * dimension receive the sum over it. has_volume_derivatives must be 0. */
dvr_status dvr_forward_variables_to_gradients(const dvr_tensor* forward_variables, const dvr_tensor* adj_color,
                                              const dvr_forward_differences* settings,
                                              dvr_adjoint_outputs* adj_outputs, void* stream);

/* Renderer::compareToImage (renderer/renderer.h:220-248; renderer/renderer.cpp:562-658): per pixel the
 * Euclidean RGBA distance L = |color - reference| and its forward derivatives
 * dot(color - reference, gradients[..., d, :]) / L (0 where L == 0).
 * reduce != 0: difference_out (1) and gradients_out (D) receive the mean over
 * all B*H*W pixels; else difference_out is B x H x W x 1 and gradients_out
 * B x H x W x D.  Any D (the reference instantiates D == 1 only). */
dvr_status dvr_compare_to_image(const dvr_tensor* color, const dvr_tensor* gradients, const dvr_tensor* reference,
                                dvr_tensor* difference_out, dvr_tensor* gradients_out, int32_t reduce, void* stream);

/* ---- volume files, synthetic volumes, TF builders (host memory) ----------
 * Voxel buffers are dense float32 host arrays with x fastest: index
 * x + X*(y + Y*z), the layout of renderer::Volume's blob (renderer/volume.cpp:21-33). */

#define DVR_CVOL_UCHAR 0  /* normalised by 1/255 on read */
#define DVR_CVOL_USHORT 1 /* normalised by 1/65535 on read */
#define DVR_CVOL_FLOAT 2

/* The 64-byte .cvol header (renderer/volume.cpp:75-84). */
typedef struct dvr_cvol_header {
    int64_t size[3];       /* X, Y, Z */
    double voxel_size[3];  /* world size = voxel_size * size */
    int32_t type;          /* DVR_CVOL_* */
    int32_t compressed;    /* LZ4 chunks follow the header */
} dvr_cvol_header;

/* Volume::Volume(filename) (renderer/volume.cpp:160-285), split into header and voxels.
 * LZ4-compressed files return DVR_ERR_UNSUPPORTED. */
dvr_status dvr_cvol_read_header(const char* path, dvr_cvol_header* header);
dvr_status dvr_cvol_read(const char* path, float* data, int64_t count);

/* Volume::save (renderer/volume.cpp:86-158): float voxels, voxel size = world size / size.
 * Only compression 0 (uncompressed) is written; 1..9 return DVR_ERR_UNSUPPORTED. */
dvr_status dvr_cvol_write(const char* path, const float* data, const int64_t size[3], const double world_size[3],
                          int32_t compression);

/* Volume::ImplicitEquation (renderer/volume.h:163-181, bindings/bindings.cpp:234-249). */
#define DVR_IMPLICIT_MARSCHNER_LOBB 0 /* parameters "fM", "alpha" */
#define DVR_IMPLICIT_CUBE 1           /* parameter "scale" */
#define DVR_IMPLICIT_SPHERE 2
#define DVR_IMPLICIT_INVERSE_SPHERE 3
#define DVR_IMPLICIT_DING_DONG 4
#define DVR_IMPLICIT_ENDRASS 5
#define DVR_IMPLICIT_BARTH 6
#define DVR_IMPLICIT_HEART 7
#define DVR_IMPLICIT_KLEINE 8
#define DVR_IMPLICIT_CASSINI 9        /* parameter "a" */
#define DVR_IMPLICIT_STEINER 10
#define DVR_IMPLICIT_CROSS_CAP 11
#define DVR_IMPLICIT_KUMMER 12
#define DVR_IMPLICIT_BLOBBY 13
#define DVR_IMPLICIT_TUBE 14

/* Volume::createImplicitDataset (renderer/volume_implicit.cpp:218-236, renderer/volume.cpp:369-387):
 * resolution^3 voxels sampled on the equation's box; named float parameters. */
dvr_status dvr_volume_implicit(int32_t equation, int32_t resolution, int32_t nparams, const char* const* names,
                               const float* values, float* data);

/* Volume::extractHistogram (renderer/volume.cpp:287-356): 512 bins over the non-zero voxels
 * (fractions summing to 1), density minimum / maximum, non-zero count.  Any element strides. */
#define DVR_HISTOGRAM_BINS 512
dvr_status dvr_volume_histogram(const float* data, const int64_t size[3], const int64_t strides[3], float* bins,
                                float* min_density, float* max_density, uint32_t* nonzero);

/* renderer::TFPoint (renderer/tf_utils.h:88-92): position and RGB + opacity. */
typedef struct dvr_tf_point {
    float pos;
    float val[4];
} dvr_tf_point;

/* sRGB <-> CIE L*a*b* (renderer/tf_utils.h:20-86). */
void dvr_rgb_to_lab(const float rgb[3], float lab[3]);
void dvr_lab_to_rgb(const float lab[3], float rgb[3]);

/* TFUtils::assembleFromSettings (renderer/tf_utils.cpp:12-118): merge the editor's colour axis (L*a*b*)
 * and opacity axis into control points.  *count receives the number of points; points (may be NULL)
 * receives them when capacity suffices. */
dvr_status dvr_tf_assemble_from_settings(const float* color_lab, const float* color_pos, int32_t nc,
                                         const float* opacity, const float* opacity_pos, int32_t no,
                                         float min_density, float max_density, float opacity_scaling,
                                         int32_t purge_zero_regions, dvr_tf_point* points, int32_t capacity,
                                         int32_t* count);

/* TFUtils::getTextureTensor (renderer/tf_utils.cpp:137-172): resolution x 4 texture sampled at the
 * texel centres. */
dvr_status dvr_tf_texture_from_points(const dvr_tf_point* points, int32_t n, int32_t resolution, float* texture);

/* ---- camera (replaces renderer/camera_impl.cu:68-93 and :302-357) ------- */

/* Camera::generateRays forward: viewport B x 3 x 3 (eye, right, up) -> per-pixel
 * ray_start, ray_dir (B x H x W x 3). */
dvr_status dvr_generate_rays(const dvr_tensor* viewport, float fov_y_radians, int32_t width, int32_t height,
                             dvr_tensor* ray_start, dvr_tensor* ray_dir, void* stream);

/* GenerateRayFunction::backward: adj_viewport (B x 3 x 3) is OVERWRITTEN with
 * the per-camera sum of the per-pixel adjoints. */
dvr_status dvr_generate_rays_adjoint(const dvr_tensor* viewport, float fov_y_radians, int32_t width,
                                     int32_t height, const dvr_tensor* adj_ray_start,
                                     const dvr_tensor* adj_ray_dir, dvr_tensor* adj_viewport, void* stream);

/* ---- batched persistent rendering ---------------------------------------
 *
 * Many views (cameras) of many volumes and transfer functions, of uneven image
 * size and resolution, rendered forward, compared with a target image and
 * differentiated in ONE persistent launch.  The reference renders a batch
 * only when all views share one image size and volume shape
 * (renderer/renderer.cpp:48-391): a caller with uneven views loops over
 * Renderer::renderForward / renderAdjoint per view
 * (pytests/diffdvr/renderer.py:127-213) and sums the gradients in torch.
 *
 * A plan binds device buffers once (set-up: it allocates and uploads the
 * descriptor tables and checkpoint scratch); dvr_batch_plan_run then only
 * launches kernels on the given stream -- no allocation, no host copy, no
 * memset node -- so a caller may capture optimisation steps into a HIP graph.
 * The plan keeps POINTERS: volume, TF and camera data are read at every run,
 * so in-place parameter updates are seen by the next run.  All buffers must
 * outlive the plan. */

typedef struct dvr_batch_volume {
    const float* data;       /* X x Y x Z contiguous (z fastest), device memory */
    int32_t size_x, size_y, size_z;
    float box_min[3], box_size[3];
    float* adj_data;         /* X x Y x Z gradient, accumulated; NULL = none */
} dvr_batch_volume;

typedef struct dvr_batch_tf {
    const float* data;       /* R x C contiguous, device memory */
    int32_t rows, cols;
    float* adj_data;         /* R x C gradient, accumulated; NULL = none */
} dvr_batch_tf;

typedef struct dvr_batch_view {
    int32_t volume, tf;      /* indices into the volume / TF tables */
    int32_t width, height;
    int32_t camera_mode;     /* DVR_CAMERA_REFERENCE_FRAME or DVR_CAMERA_INVERSE_VIEW_MATRIX */
    const float* camera;     /* device: 3x3 viewport rows (eye, right, up) or 4x4 row-major matrix */
    float fov_y_radians;     /* ReferenceFrame */
    float step_size;
    float* color;            /* H x W x 4 */
    int32_t* termination_index; /* H x W */
    const float* target;     /* H x W x 4 loss target (loss kind L1/L2) */
    const float* adj_color;  /* H x W x 4 given adjoint of the colour (loss kind none) */
    float loss_scale;        /* factor of this view's summed loss; 0 = 1/(4 * all pixels of the batch) */
    float* adj_camera;       /* 3 x 3 viewport gradient (ReferenceFrame), accumulated; NULL = none */
} dvr_batch_view;

#define DVR_LOSS_NONE 0      /* adjoint colour given per view */

typedef struct dvr_batch {
    int32_t num_volumes, num_tfs, num_views;
    const dvr_batch_volume* volumes;  /* host arrays, copied at plan creation */
    const dvr_batch_tf* tfs;
    const dvr_batch_view* views;
    int32_t volume_filter_mode, tf_mode, blend_mode;
    float blending_early_out;         /* 0 = 1 - 1e-5 */
    int32_t loss;                     /* DVR_LOSS_NONE / L1 / L2 */
    float* loss_out;                  /* device scalar, the loss is ADDED to it; NULL = none */
    int32_t forward_only;             /* 1 = colour and termination only */
    int32_t device;                   /* HIP device ordinal of every buffer */
} dvr_batch;

typedef struct dvr_batch_options {
    int32_t checkpoint_interval; /* K in {4,8,16,32,64}; 0 = 16 */
    int32_t wave_width;          /* rays along x of a wave's tile: 4, 8, 16 or 32 (tile 4x16 .. 32x2); 0 = 8 */
    int32_t grid;                /* persistent workgroups; 0 = 2 per compute unit (within residency) */
    int32_t reinterpolate;       /* 0: the forward march stores every sample's density and the reverse march
                                    reads it back (when the store fits in 2 GiB); 1: the reverse march
                                    re-interpolates (no store) */
} dvr_batch_options;

typedef struct dvr_batch_plan dvr_batch_plan;

#define DVR_BATCH_ZERO_GRADIENTS 1   /* run: clear every gradient buffer and loss_out first */

dvr_status dvr_batch_plan_create(const dvr_batch* batch, const dvr_batch_options* options, dvr_batch_plan** plan);
dvr_status dvr_batch_plan_destroy(dvr_batch_plan* plan);
/* One persistent launch (plus one clearing kernel) on `stream`. */
dvr_status dvr_batch_plan_run(dvr_batch_plan* plan, int32_t flags, void* stream);
/* Clear the gradient buffers and loss_out with a kernel (no memset). */
dvr_status dvr_batch_plan_zero_gradients(dvr_batch_plan* plan, void* stream);
/* Measurement: one run (as dvr_batch_plan_run) with HIP events recorded on `stream` around the
 * persistent kernel alone; waits for it and writes that kernel's duration.  No counterpart in the
 * reference (its GPUTimer, bindings/bindings.cpp:99-129, times whatever the caller brackets). */
dvr_status dvr_batch_plan_time_kernel(dvr_batch_plan* plan, int32_t flags, void* stream, float* kernel_ms);
/* Freeze the plan's layout: from now on set_options may only repeat the current options and
 * autotune fails.  Call before capturing runs into a HIP graph: a captured graph holds the
 * plan's scratch pointers and kernel arguments, which a re-layout would free or change. */
dvr_status dvr_batch_plan_freeze(dvr_batch_plan* plan);
dvr_status dvr_batch_plan_get_options(const dvr_batch_plan* plan, dvr_batch_options* options);
/* Re-layout for new options; may allocate (set-up only, not under capture). */
dvr_status dvr_batch_plan_set_options(dvr_batch_plan* plan, const dvr_batch_options* options);
/* Autotuner: times candidates on this plan's own batch (`iterations` timed
 * runs each, HIP events) in two stages -- checkpoint interval {4,8,16,32,64}
 * jointly with workgroups per CU (1..4 waves per SIMD and full residency, per
 * interval, unless options.grid was set), then wave tile width {8,16,4,32} --
 * keeps the fastest and writes it to *chosen.  Results are
 * cached per device and batch signature (GPU architecture and compute units,
 * modes, gradients, every volume / TF / view shape and step) for the process
 * and in a text file: $DVR_AUTOTUNE_CACHE, else
 * $XDG_CACHE_HOME/diffdvr_amd/autotune.txt, else
 * $HOME/.cache/diffdvr_amd/autotune.txt (DVR_AUTOTUNE_CACHE="" turns the file
 * off).  A cached result is applied without timing runs.  Timing runs
 * overwrite the outputs; the gradients are left ZEROED either way. */
dvr_status dvr_batch_plan_autotune(dvr_batch_plan* plan, int32_t iterations, dvr_batch_options* chosen, void* stream);
/* Slots of the per-wave LDS table that sums the volume-gradient corners per 64-byte chunk
 * before the device atomics (the largest of 32..256 that keeps the volume-gradient instance at
 * four waves per SIMD beside the plan's TF tables, for the current options); 0 when the plan
 * takes no volume gradients, -1 for NULL.  No counterpart in the reference (it scatters every
 * corner with a device atomic, renderer_interpolation.cuh:299-324). */
int32_t dvr_batch_plan_chunk_slots(const dvr_batch_plan* plan);
/* Bytes of device scratch the plan holds (checkpoints, tables). */
int64_t dvr_batch_plan_workspace_bytes(const dvr_batch_plan* plan);

/* ---- optimiser step ------------------------------------------------------
 * Adam (torch.optim.Adam, no weight decay, no amsgrad; the reference's
 * optimisation scripts use it, pytests/tests/volume/compare_reconstruction.py:64,414)
 * over several parameter blocks in one launch.  Block b is rows x cols floats
 * with row stride `stride` shared by param, grad, exp_avg and exp_avg_sq;
 * after the update the parameter is clamped to [clamp_min, clamp_max] when
 * clamp_min <= clamp_max. */
typedef struct dvr_adam_block {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t rows, cols, stride;
    float clamp_min, clamp_max;
} dvr_adam_block;

typedef struct dvr_adam_group dvr_adam_group;
/* Uploads the block table once (set-up). */
dvr_status dvr_adam_group_create(const dvr_adam_block* blocks, int32_t num_blocks, int32_t device,
                                 dvr_adam_group** group);
dvr_status dvr_adam_group_destroy(dvr_adam_group* group);
/* One Adam step: advances the group's device-side step counter t (0 at
 * creation) and updates every block with bias corrections 1 - beta^t.  Two
 * launches, no allocation, no host copy: captured steps keep counting across
 * graph replays.  lr, betas and eps are given in double as torch's Python
 * floats and rounded to float where torch rounds them. */
dvr_status dvr_adam_group_step(dvr_adam_group* group, double lr, double beta1, double beta2, double eps,
                               void* stream);
/* Set-up only (synchronous copies): set or read the step counter. */
dvr_status dvr_adam_group_set_step(dvr_adam_group* group, int64_t step);
dvr_status dvr_adam_group_get_step(dvr_adam_group* group, int64_t* step);

/* ---- TF utilities on device tensors ---------------------------------------
 * Replace renderer::TFUtils::preshadeVolume / findBestFit
 * (renderer/tf_utils.cpp:173-334, bound at bindings/bindings.cpp:570-633)
 * and kernel::PreshadeVolume(Adj) / FindBestFit(WithComparison)
 * (renderer/tf_utils_impl.cu:56-190, :301-520).  Volumes are 1 x X x Y x Z
 * (densities) or 4 x X x Y x Z (r, g, b, absorption), float32, any strides;
 * tf is 1 x R x C as for rendering: identity 1 x 2, texture R x 4, linear
 * R x 5 (R > 1), Gaussian and GaussianLog R x 6.  R*C <= 16384. */

/* color = TF(density) per voxel (tf_mode identity/texture/linear/Gaussian/GaussianLog). */
dvr_status dvr_preshade_volume(const dvr_tensor* density, const dvr_tensor* tf, int32_t tf_mode, dvr_tensor* color,
                               void* stream);
/* adj_density = d<adj_color, TF(density)>/d density per voxel (overwritten, not accumulated). */
dvr_status dvr_preshade_volume_adjoint(const dvr_tensor* density, const dvr_tensor* tf, int32_t tf_mode,
                                       const dvr_tensor* adj_color, dvr_tensor* adj_density, void* stream);
/* density = the best of num_samples random densities per voxel for the target colours
 * (tf_mode identity/texture/linear/Gaussian).  previous_densities (NULL or not given =
 * none) adds neighbor_weighting * (d - previous neighbour)^2 for each existing axis
 * neighbour.  Each voxel draws its own candidates from `seed` and its (x, y, z) index,
 * so the result does not depend on the launch (the reference's draws do). */
dvr_status dvr_find_best_fit(const dvr_tensor* color_volume, const dvr_tensor* tf, int32_t tf_mode,
                             int32_t num_samples, float opacity_weighting, const dvr_tensor* previous_densities,
                             float neighbor_weighting, uint64_t seed, dvr_tensor* density, void* stream);

/* ---- importance sampling and elementwise losses ---------------------------
 * Replace renderer::sampleImportance / mul_log / logMSE / logL1
 * (renderer/losses.cpp:8-221, bound at bindings/bindings.cpp:636-688) and
 * their kernels (renderer/losses_impl.cu:35-421). */

/* out[i] = whether sample i (sample_locations N x 3 in [0,1]^3, float32) is taken, as
 * DVR_BOOL (N).  weights: uniform, density gradient, opacity, opacity gradient (>= 0,
 * normalised here).  maxima: float32 tensor of 3 elements, overwritten with the maxima
 * of |grad density|, opacity and |grad opacity| over the voxels (computed on the
 * device; the reference copies them to the host).  Sample i draws its two uniform
 * numbers from `seed` and i. */
dvr_status dvr_sample_importance(const dvr_tensor* sample_locations, const dvr_tensor* density, const dvr_tensor* tf,
                                 int32_t tf_mode, const float weights[4], int32_t seed, dvr_tensor* maxima,
                                 dvr_tensor* out, void* stream);
/* Elementwise over contiguous tensors of one shape and dtype (float32 or float64). */
dvr_status dvr_mul_log(const dvr_tensor* t, dvr_tensor* out, void* stream);
dvr_status dvr_mul_log_backward(const dvr_tensor* t, const dvr_tensor* grad_out, dvr_tensor* grad_t, void* stream);
dvr_status dvr_log_mse(const dvr_tensor* log_x, const dvr_tensor* log_y, dvr_tensor* out, void* stream);
dvr_status dvr_log_mse_backward(const dvr_tensor* log_x, const dvr_tensor* log_y, const dvr_tensor* grad_out,
                                dvr_tensor* grad_log_x, dvr_tensor* grad_log_y, void* stream);
dvr_status dvr_log_l1(const dvr_tensor* log_x, const dvr_tensor* log_y, dvr_tensor* out, void* stream);
dvr_status dvr_log_l1_backward(const dvr_tensor* log_x, const dvr_tensor* log_y, const dvr_tensor* grad_out,
                               dvr_tensor* grad_log_x, dvr_tensor* grad_log_y, void* stream);

/* ---- GPU timer (bindings/bindings.cpp:99-129) ---------------------------- */
typedef struct dvr_timer dvr_timer;
dvr_status dvr_timer_create(dvr_timer** timer);
dvr_status dvr_timer_destroy(dvr_timer* timer);
dvr_status dvr_timer_start(dvr_timer* timer, void* stream);
dvr_status dvr_timer_stop(dvr_timer* timer, void* stream);
dvr_status dvr_timer_elapsed_ms(dvr_timer* timer, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* DVR_H_ */
