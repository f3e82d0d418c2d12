/*
 * swr.h -- C ABI of the MI355X fixed-function triangle draw path.
 *
 * One call renders one triangle-list draw into caller-owned device memory,
 * with the semantics of SwiftShader's sw::Renderer::draw for the FILL
 * polygon mode: post-vertex-stage clip-space vertices in, colour / depth /
 * stencil attachments and an occlusion count out.
 *
 * Reference interface replaced (SwiftShader, legacy 28.4 precision build):
 *   sw::Renderer::draw(const vk::GraphicsPipeline*, const vk::DynamicState&,
 *       unsigned count, int baseVertex, CountedEvent*, int instanceID,
 *       int layer, void *indexBuffer, const VkRect2D &renderArea,
 *       const vk::Pipeline::PushConstantStorage&, bool update)
 *                                             src/Device/Renderer.hpp:211-213
 *   completion: sw::Renderer::synchronize()    src/Device/Renderer.hpp:218
 *   MSAA resolve: sw::Blitter::resolve()       src/Device/Blitter.hpp:147
 *
 * The descriptor carries what DrawCall / DrawData carry
 * (src/Device/Renderer.hpp:58-113): vertex outputs, viewport and scissor,
 * the SetupProcessor / PixelProcessor state subsets, blend constants,
 * stencil reference and masks, the sw::Sampler subset and per-level mip
 * descriptors, and the attachment pointers, pitches and sample slices.
 *
 * Every pointer inside a descriptor is a DEVICE pointer (HBM), owned by the
 * caller.  The library owns only its scratch (setup records, tile bins),
 * grown on demand.  All enum values are the Vulkan ones, so a Vulkan front
 * end can copy them through unchanged.
 *
 * Error behaviour: functions return SWR_OK (0) or a negative SWR_ERROR_*
 * code; they never abort.  swr_last_error() returns a message for the last
 * failure on the calling thread.  A primitive that the reference would cull
 * (trivially rejected, clipped away, back-facing, zero coverage) is simply
 * not drawn -- that is not an error, as in SetupRoutine.
 *
 * Threading: calls are thread-compatible; one stream per caller thread.
 * Work is enqueued on the given HIP stream (NULL = default stream) and is
 * complete when the stream is synchronised.
 */
#ifndef SWR_H_
#define SWR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWR_ABI_VERSION 12

/* ---- status codes -------------------------------------------------------- */
#define SWR_OK 0
#define SWR_ERROR_INVALID_ARGUMENT (-1)
#define SWR_ERROR_UNSUPPORTED (-2)
#define SWR_ERROR_OUT_OF_MEMORY (-3)
#define SWR_ERROR_DEVICE (-4)

/* ---- Vulkan enum values used by the descriptor -------------------------- */
/* VkFormat */
#define SWR_FORMAT_UNDEFINED 0
#define SWR_FORMAT_R8G8B8A8_UNORM 37
#define SWR_FORMAT_R8G8B8A8_SRGB 43
#define SWR_FORMAT_B8G8R8A8_UNORM 44
#define SWR_FORMAT_B8G8R8A8_SRGB 50
#define SWR_FORMAT_R32_UINT 98
#define SWR_FORMAT_R32G32B32A32_SFLOAT 109
#define SWR_FORMAT_D16_UNORM 124
#define SWR_FORMAT_D32_SFLOAT 126
#define SWR_FORMAT_S8_UINT 127
/* VkCompareOp */
#define SWR_COMPARE_NEVER 0
#define SWR_COMPARE_LESS 1
#define SWR_COMPARE_EQUAL 2
#define SWR_COMPARE_LESS_OR_EQUAL 3
#define SWR_COMPARE_GREATER 4
#define SWR_COMPARE_NOT_EQUAL 5
#define SWR_COMPARE_GREATER_OR_EQUAL 6
#define SWR_COMPARE_ALWAYS 7
/* VkStencilOp */
#define SWR_STENCIL_OP_KEEP 0
#define SWR_STENCIL_OP_ZERO 1
#define SWR_STENCIL_OP_REPLACE 2
#define SWR_STENCIL_OP_INCREMENT_AND_CLAMP 3
#define SWR_STENCIL_OP_DECREMENT_AND_CLAMP 4
#define SWR_STENCIL_OP_INVERT 5
#define SWR_STENCIL_OP_INCREMENT_AND_WRAP 6
#define SWR_STENCIL_OP_DECREMENT_AND_WRAP 7
/* VkBlendFactor */
#define SWR_BLEND_FACTOR_ZERO 0
#define SWR_BLEND_FACTOR_ONE 1
#define SWR_BLEND_FACTOR_SRC_COLOR 2
#define SWR_BLEND_FACTOR_ONE_MINUS_SRC_COLOR 3
#define SWR_BLEND_FACTOR_DST_COLOR 4
#define SWR_BLEND_FACTOR_ONE_MINUS_DST_COLOR 5
#define SWR_BLEND_FACTOR_SRC_ALPHA 6
#define SWR_BLEND_FACTOR_ONE_MINUS_SRC_ALPHA 7
#define SWR_BLEND_FACTOR_DST_ALPHA 8
#define SWR_BLEND_FACTOR_ONE_MINUS_DST_ALPHA 9
#define SWR_BLEND_FACTOR_CONSTANT_COLOR 10
#define SWR_BLEND_FACTOR_ONE_MINUS_CONSTANT_COLOR 11
#define SWR_BLEND_FACTOR_CONSTANT_ALPHA 12
#define SWR_BLEND_FACTOR_ONE_MINUS_CONSTANT_ALPHA 13
#define SWR_BLEND_FACTOR_SRC_ALPHA_SATURATE 14
/* VkBlendOp */
#define SWR_BLEND_OP_ADD 0
#define SWR_BLEND_OP_SUBTRACT 1
#define SWR_BLEND_OP_REVERSE_SUBTRACT 2
#define SWR_BLEND_OP_MIN 3
#define SWR_BLEND_OP_MAX 4
/* VK_EXT_blend_operation_advanced, the modes the reference implements
 * (PixelRoutine.cpp:1506-1615); the colour and alpha ops must be equal. */
#define SWR_BLEND_OP_MULTIPLY_EXT 1000148012
#define SWR_BLEND_OP_SCREEN_EXT 1000148013
#define SWR_BLEND_OP_OVERLAY_EXT 1000148014
#define SWR_BLEND_OP_DARKEN_EXT 1000148015
#define SWR_BLEND_OP_LIGHTEN_EXT 1000148016
#define SWR_BLEND_OP_COLORDODGE_EXT 1000148017
#define SWR_BLEND_OP_COLORBURN_EXT 1000148018
#define SWR_BLEND_OP_HARDLIGHT_EXT 1000148019
#define SWR_BLEND_OP_SOFTLIGHT_EXT 1000148020
#define SWR_BLEND_OP_DIFFERENCE_EXT 1000148021
#define SWR_BLEND_OP_EXCLUSION_EXT 1000148022
#define SWR_BLEND_OP_HSL_HUE_EXT 1000148031
#define SWR_BLEND_OP_HSL_SATURATION_EXT 1000148032
#define SWR_BLEND_OP_HSL_COLOR_EXT 1000148033
#define SWR_BLEND_OP_HSL_LUMINOSITY_EXT 1000148034
/* VkPrimitiveTopology / VkProvokingVertexModeEXT */
#define SWR_TOPOLOGY_POINT_LIST 0
#define SWR_TOPOLOGY_LINE_LIST 1
#define SWR_TOPOLOGY_LINE_STRIP 2
#define SWR_TOPOLOGY_TRIANGLE_LIST 3
#define SWR_TOPOLOGY_TRIANGLE_STRIP 4
#define SWR_TOPOLOGY_TRIANGLE_FAN 5
#define SWR_LINE_RASTERIZATION_DEFAULT 0 /* VkLineRasterizationModeEXT */
#define SWR_LINE_RASTERIZATION_RECTANGULAR 1
#define SWR_LINE_RASTERIZATION_BRESENHAM 2
#define SWR_LINE_RASTERIZATION_RECTANGULAR_SMOOTH 3
#define SWR_INDEX_TYPE_UINT16 0 /* VkIndexType */
#define SWR_INDEX_TYPE_UINT32 1
#define SWR_INDEX_TYPE_UINT8 1000265000
#define SWR_PROVOKING_VERTEX_FIRST 0
#define SWR_PROVOKING_VERTEX_LAST 1
/* VkCullModeFlagBits / VkFrontFace */
#define SWR_CULL_NONE 0
#define SWR_CULL_FRONT 1
#define SWR_CULL_BACK 2
#define SWR_CULL_FRONT_AND_BACK 3
#define SWR_FRONT_FACE_COUNTER_CLOCKWISE 0
#define SWR_FRONT_FACE_CLOCKWISE 1
/* VkFilter / VkSamplerMipmapMode / VkSamplerAddressMode */
#define SWR_FILTER_NEAREST 0
#define SWR_FILTER_LINEAR 1
#define SWR_MIPMAP_MODE_NEAREST 0
#define SWR_MIPMAP_MODE_LINEAR 1
#define SWR_ADDRESS_REPEAT 0
#define SWR_ADDRESS_MIRRORED_REPEAT 1
#define SWR_ADDRESS_CLAMP_TO_EDGE 2
#define SWR_ADDRESS_MIRROR_CLAMP_TO_EDGE 4

/* ---- fixed fragment-shader family ---------------------------------------- */
/* Inputs: location 0 = vec4 colour, location 1 = vec2 texcoord, both
 * perspective-correct.  Output: location 0 vec4. */
#define SWR_SHADER_INTERPOLATE 0 /* out = colour                           */
#define SWR_SHADER_SAMPLE 1      /* out = texture(s, uv)                    */
#define SWR_SHADER_MODULATE 2    /* out = texture(s, uv) * colour           */

#define SWR_MAX_MIP_LEVELS 15 /* sw::MIPMAP_LEVELS, src/Device/Config.hpp:21 */

typedef struct swr_viewport { /* VkViewport */
  float x, y, width, height, min_depth, max_depth;
} swr_viewport;

typedef struct swr_rect { /* VkRect2D */
  int32_t x, y;
  uint32_t width, height;
} swr_rect;

typedef struct swr_stencil_op_state { /* VkStencilOpState */
  uint32_t fail_op, pass_op, depth_fail_op, compare_op;
  uint32_t compare_mask, write_mask, reference;
} swr_stencil_op_state;

typedef struct swr_blend_state { /* VkPipelineColorBlendAttachmentState */
  uint32_t blend_enable;
  uint32_t src_color_blend_factor, dst_color_blend_factor, color_blend_op;
  uint32_t src_alpha_blend_factor, dst_alpha_blend_factor, alpha_blend_op;
  uint32_t color_write_mask; /* bit0 R .. bit3 A */
} swr_blend_state;

typedef struct swr_sampler_state { /* VkSamplerCreateInfo subset */
  uint32_t mag_filter, min_filter, mipmap_mode;
  uint32_t address_mode_u, address_mode_v;
  float mip_lod_bias;
  uint32_t anisotropy_enable;
  float max_anisotropy;
  float min_lod, max_lod; /* clamped to [0, 13] as vk::SamplerState does */
} swr_sampler_state;

typedef struct swr_image_level { /* one mip level of a 2D image view */
  const void *data;         /* device pointer to texel (0,0)            */
  uint32_t width, height;   /* texels                                   */
  uint32_t row_pitch_bytes; /* multiple of the texel size               */
  uint32_t reserved;
} swr_image_level;

typedef struct swr_texture { /* combined image sampler, 2D view */
  uint32_t format;          /* R8G8B8A8 / B8G8R8A8 (UNORM or SRGB), or R32G32B32A32_SFLOAT */
  uint32_t level_count;     /* levels in the view, 1..15                */
  swr_image_level levels[SWR_MAX_MIP_LEVELS];
} swr_texture;

typedef struct swr_attachment {
  void *data;                  /* device pointer to sample 0, pixel (0,0) */
  uint32_t format;             /* colour: RGBA8 / BGRA8 (UNORM, sRGB) / RGBA32F; depth: D32F / D16; stencil: S8 */
  uint32_t row_pitch_bytes;
  uint64_t sample_pitch_bytes; /* distance between sample planes (4x MSAA) */
} swr_attachment;

typedef struct swr_draw_desc {
  /* Vertex-stage outputs, one entry per vertex (structure of arrays). */
  const float *positions;   /* float4 clip-space x, y, z, w (required) */
  const float *colors;      /* float4 (shader input location 0) or NULL */
  const float *texcoords;   /* float2 (shader input location 1) or NULL */
  uint32_t vertex_count;
  uint32_t triangle_count;  /* primitives (triangles, or lines for the line topologies) */
  const void *indices;      /* of index_type, or NULL for 0,1,2,3,...   */

  /* Pre-rasterisation state. */
  swr_viewport viewport;
  swr_rect scissor;
  swr_rect render_area;
  uint32_t cull_mode;
  uint32_t front_face;

  /* Multisample state. */
  uint32_t samples;      /* 1 or 4 */
  uint32_t sample_mask;  /* VkPipelineMultisampleStateCreateInfo.pSampleMask[0] */

  /* Depth / stencil state. */
  uint32_t depth_test_enable, depth_write_enable, depth_compare_op;
  uint32_t stencil_test_enable;
  swr_stencil_op_state front, back;

  /* Colour blend state (one attachment). */
  swr_blend_state blend;
  float blend_constants[4];

  /* Fragment shader (fixed family) and its combined image sampler. */
  uint32_t shader;
  swr_sampler_state sampler;
  swr_texture texture;

  /* Attachments; a NULL data pointer means "no such attachment". */
  uint32_t framebuffer_width, framebuffer_height;
  swr_attachment color, depth, stencil;

  /* Optional device pointer: += samples passing depth/stencil
   * (the VK_QUERY_TYPE_OCCLUSION result of the reference). */
  uint64_t *occlusion_count;

  /* Optional R32_UINT attachment, one plane per sample (same pitches as the
   * other attachments): each sample whose colour a fragment writes receives
   * 1 + that fragment's primitive index in draw order (0 = never written).
   * This is the visibility record swr_draw_backward reads.  It has no
   * counterpart in the reference; it requires an RGBA32F colour attachment. */
  swr_attachment primitive_id;

  /* Primitive assembly (setBatchIndices, src/Device/Renderer.cpp:50-145):
   * the VkPrimitiveTopology -- POINT_LIST (0), LINE_LIST (1), LINE_STRIP
   * (2), TRIANGLE_LIST (3), TRIANGLE_STRIP (4) or TRIANGLE_FAN (5) -- over
   * the index buffer (or the vertex sequence when indices is NULL), and the
   * VkProvokingVertexModeEXT (0 first, 1 last), which rotates each
   * primitive's vertices.  triangle_count counts primitives (points or lines
   * for those topologies): a triangle list needs 3 * triangle_count indices,
   * a line list 2 * triangle_count, a strip or fan triangle_count + 2, a
   * line strip triangle_count + 1, a point list triangle_count. */
  uint32_t topology;
  uint32_t provoking_vertex;

  /* Depth bias (VkPipelineRasterizationStateCreateInfo): the setup adds
   * r * constant + max(|dz/dx|, |dz/dy|) * slope, clamped by depth_bias_clamp
   * when it is non-zero (a NaN clamp is no clamp), to every sample's depth
   * before the depth clamp and test (SetupRoutine.cpp:417-475,
   * PixelRoutine.cpp:141-144, Context.cpp:640-661).  D32F: r = 2^(e - 23),
   * e the largest exponent of the depth plane's constant and of
   * (z1 - z0) * range + near, (z2 - z0) * range + near, as the reference
   * forms them. */
  uint32_t depth_bias_enable;
  float depth_bias_constant_factor;
  float depth_bias_clamp;
  float depth_bias_slope_factor;

  /* Lines (DrawCall::setupLine, Renderer.cpp:920-1163).  line_width is
   * lineWidth (0 reads as 1; accepted in (0, 64] -- the reference device
   * exposes only 1, VkPhysicalDevice.cpp:53,820, and builds wider lines the
   * same way).  line_rasterization_mode is the
   * VkLineRasterizationModeEXT: BRESENHAM (2) sets up the parallelogram
   * approximating a Bresenham line and needs samples == 1 (the reference
   * turns multisampling off for it); every other mode sets up the
   * rectangle of width line_width centred on the segment.  Either shape is
   * clipped against the frustum and rasterised as a polygon.  Depth bias
   * applies to triangles only and is refused for lines. */
  float line_width;
  uint32_t line_rasterization_mode;

  /* Points (DrawCall::setupPoint, Renderer.cpp:1166-1210): each point is
   * the square of side PointSize -- the vertex's point_sizes entry (the
   * vertex stage's PointSize output), 1 when point_sizes is NULL, clamped
   * to [1, 1023] -- centred on the vertex, clipped and rasterised as a
   * polygon with a constant depth.  Depth bias is refused for points. */
  const float *point_sizes; /* one per vertex, optional */

  /* VkIndexType of `indices` (setBatchIndices is instantiated for each,
   * Renderer.cpp:696-716): UINT16 (0), UINT32 (1) or UINT8_EXT
   * (1000265000).  Primitive restart is the caller's: the reference splits
   * the index buffer into separate draws before Renderer::draw
   * (Context.cpp:48-90, 249-285).  A zero-initialised descriptor reads
   * 16-bit indices: callers written for ABI 5 or earlier, whose index
   * buffers were always 32-bit, must set SWR_INDEX_TYPE_UINT32. */
  uint32_t index_type;

  /* Depth clamp and depth clip (Context.cpp:647-648, 686-693;
   * PixelProcessor.cpp:120-136; VertexRoutine.cpp:137-143; Renderer.cpp:940,
   * 1176).  depth_clamp_enable clamps every sample's depth to the viewport's
   * [min, max] depth range instead of [0, 1].  depth_clip_control is the
   * VK_EXT_depth_clip_enable state: 0 uses the default (clip unless
   * depth_clamp_enable), 1 clips, 2 does not clip against the near and far
   * planes. */
  uint32_t depth_clamp_enable;
  uint32_t depth_clip_control;

  /* Depth bounds test (PixelRoutine.cpp:576-641; Context.cpp:946-962).  With
   * a depth attachment, a sample whose stored depth (D16: code / 0xFFFF) lies
   * outside [min_depth_bounds, max_depth_bounds] loses its depth-pass bit when
   * the depth test is on (no depth or colour write; the stencil takes its
   * depth-fail op) and its coverage when it is off (no colour or stencil
   * write; the occlusion count still includes it). */
  uint32_t depth_bounds_test_enable;
  float min_depth_bounds;
  float max_depth_bounds;

  /* Alpha-to-coverage (VkPipelineMultisampleStateCreateInfo::
   * alphaToCoverageEnable; PixelProgram.cpp:243-259, PixelRoutine.cpp:
   * 317-337, 643-657, thresholds Renderer.cpp:391-411): the fragment is
   * shaded before the depth test and sample q keeps its coverage only where
   * the shaded alpha is not below 0.5 (1 sample) or 0.2, 0.4, 0.6, 0.8
   * (4 samples) -- the alpha clamped to [0, 1] when the colour attachment
   * is UNORM, as is, NaN included, otherwise (a NaN passes).  The samples
   * it removes take no depth, stencil, colour or occlusion effect. */
  uint32_t alpha_to_coverage_enable;
} swr_draw_desc;

typedef struct swr_resolve_desc { /* 4x -> 1x colour resolve */
  uint32_t width, height;
  uint32_t format;               /* RGBA8, BGRA8 (UNORM: byte average; sRGB: linear-space average) or RGBA32F */
  const void *src;               /* multisampled source, sample 0 plane */
  uint32_t src_row_pitch_bytes;
  uint64_t src_sample_pitch_bytes;
  uint32_t src_samples;          /* 4 */
  void *dst;
  uint32_t dst_row_pitch_bytes;
} swr_resolve_desc;

typedef union swr_clear_value { /* VkClearValue */
  float color[4];                 /* RGBA8 / BGRA8 / RGBA32F targets */
  struct {
    float depth;                  /* D32F / D16 targets */
    uint32_t stencil;             /* S8 targets */
  } depth_stencil;
} swr_clear_value;

typedef struct swr_clear_desc {
  swr_attachment target;          /* format: any attachment format swr_draw takes */
  uint32_t width, height;         /* image extent in pixels */
  uint32_t samples;               /* sample planes: 1 or 4 */
  uint32_t use_area;              /* 0: the whole image; 1: `area` (clipped to the image) */
  swr_rect area;
  swr_clear_value value;
} swr_clear_desc;

/* Gradients of a draw (differentiable rendering; no counterpart in the
 * reference).  The draw is the one a swr_draw_desc describes, rendered with
 * its primitive_id attachment; blending must be off, the colour attachment
 * RGBA32F and a sampled texture RGBA32F.
 *
 * What is differentiated: coverage and visibility are discrete and held
 * fixed (the primitive_id record); so are the texture LOD, the anisotropic
 * probe count and axis, and the mip levels, which take the forward pass's
 * values.  Per covered sample, the colour is the fixed shader evaluated at
 * the pixel centre with perspective-correct barycentrics of the unsnapped
 * projected vertices, and the texture filter is the forward's
 * (bilinear / trilinear / anisotropic, REPEAT or CLAMP_TO_EDGE).  z does not
 * reach the colour, so d_positions[.z] stays 0.  All outputs are
 * ACCUMULATED (+=) with float atomics; NULL outputs are skipped. */
typedef struct swr_grad_desc {
  const float *d_color;                /* dL/d(colour attachment): float4 per sample */
  uint32_t d_color_row_pitch_bytes;
  uint64_t d_color_sample_pitch_bytes;
  float *d_positions;                  /* float4 per vertex (x, y, z, w)        */
  float *d_colors;                     /* float4 per vertex                      */
  float *d_texcoords;                  /* float2 per vertex                      */
  void *d_texture[SWR_MAX_MIP_LEVELS]; /* float4 per texel, per level, with the
                                          texture level's row pitch              */
  float *sample_state;                 /* optional float4 per sample: the frozen
                                          (lod, anisotropy, uDelta, vDelta) used */
  uint32_t sample_state_row_pitch_bytes;
  uint64_t sample_state_sample_pitch_bytes;
} swr_grad_desc;

/* Tuning knobs of the tiled rasteriser (0 = library default). */
typedef struct swr_tuning {
  uint32_t tile_width, tile_height;    /* pixels per screen tile: 8 or 16 each */
  uint32_t raster_workgroups_per_cu;   /* persistent raster workgroups per CU;
                                          ~0u = one workgroup per tile instead */
  uint32_t reserved[5];
} swr_tuning;

/* A batch of draws prepared once and enqueued many times (no counterpart
 * in the reference, whose draws are enqueued one by one): the descriptors
 * are validated and their derived state uploaded, the tiled pipeline's
 * scratch is sized from a first setup pass (the (tile, primitive) pair
 * count, plus the given headroom), and every later run enqueues a fixed
 * sequence of kernels that allocates nothing and never waits on the host --
 * so a run can be captured into a HIP graph and replayed.  The memory the
 * descriptors point to is read at run time: vertex and texel contents and
 * attachments may change between runs, triangle counts and sizes may not.
 * No primitive is ever dropped: a run whose geometry needs more (tile,
 * primitive) pairs than planned bins the primitives whose pairs fit and
 * renders the rest, in API order after them, by an unbinned pass in which
 * each tile walks those primitives directly (as the reference re-batches
 * any count, Renderer.cpp:250-268).  That pass is slower, so such a run is
 * reported by swr_batch_plan_check; rebuilding the plan (destroy + create,
 * then capturing any graph again) restores the binned speed. */
typedef struct swr_batch_plan swr_batch_plan;

/* Returns SWR_ABI_VERSION. */
int swr_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *swr_last_error(void);

/* Selects the HIP device used by later calls on this thread and prepares
 * per-device state.  Optional: the first call on a device does it. */
int swr_init(int device);

/* Releases library scratch on the current device (and the (tile,
 * primitive) pair capacity swr_draw has learnt). */
int swr_release(void);

/* Renders one draw (sw::Renderer::draw).  Reads and writes only the memory
 * the descriptor names plus library scratch.  Asynchronous on `stream`: the
 * first draw on a device (and the first after swr_release) waits once for
 * its pair count to size the library scratch; later draws size it from the
 * counts of earlier ones, read back without waiting, and render any pairs
 * beyond it through the unbinned pass (as a batch plan does), so they
 * return without waiting on the device.  Scratch grows by allocation when a
 * count shows it too small. */
int swr_draw(const swr_draw_desc *desc, void *stream);

/* Renders one draw as swr_draw does, without allocating.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It calls no hipMalloc, hipFree or host-pinned allocation and
 * never waits on the host.  Its scratch (primitive records, the (tile,
 * primitive) pairs and their sort) is one device array of
 * SWR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and shared
 * by every call on that GPU, so calls that overlap on one GPU (other
 * streams, other threads) race.  Each primitive takes 296 bytes and each
 * 16x16 screen tile 8; the pairs get what is left, and the pairs beyond it
 * are rendered by the unbinned pass (slower, never dropped).  A draw that
 * leaves room for fewer than 65,536 pairs (about 330,000 primitives)
 * returns SWR_ERROR_UNSUPPORTED.  swr_init(device) locates the device's
 * array; otherwise the first call on that device does. */
#define SWR_NOALLOC_SCRATCH_BYTES (96u << 20)
int swr_draw_noalloc(const swr_draw_desc *desc, void *stream);

/* Renders `count` independent draws in one submission, in array order
 * (draws may target different attachments).  Equivalent to `count` calls
 * of swr_draw on the same stream. */
int swr_draw_batch(const swr_draw_desc *descs, uint32_t count, void *stream);

/* MSAA colour resolve (sw::Blitter::resolve, src/Device/Blitter.cpp:2035). */
int swr_resolve(const swr_resolve_desc *desc, void *stream);

/* Attachment clear (vk::Image::clear, src/Vulkan/VkImage.cpp:1105-1124, ->
 * sw::Blitter::clear / fastClear, src/Device/Blitter.cpp:55-326): every
 * sample of the area.  RGBA8 and BGRA8 channels are clamped to [0, 1] (a
 * NaN passes the clamp and packs as 0) and packed as (uint32_t)(255 * c +
 * 0.5f), BGRA8 with R and B exchanged; the sRGB formats through the blit
 * routine (clamp, linearToSRGB of R, G, B, Round(255 * c), a NaN becoming
 * 255); RGBA32F is written as given; D32F
 * gets the depth's bits unclamped (the format is not normalised); D16 the
 * blit routine's Round(65535 * clamp(d)) (a NaN becomes 0xFFFF,
 * Blitter.cpp:1427-1484); S8 the stencil's low byte. */
int swr_clear(const swr_clear_desc *desc, void *stream);

/* count clears with the result of count swr_clear calls, in as few kernel
 * launches as their element sizes allow (up to 24 clears per launch, the
 * areas passed as kernel arguments: nothing is uploaded, so the call can be
 * captured into a graph).  No counterpart in the reference: a render pass
 * clears its attachments one vk::Image::clear at a time; this is for
 * callers that clear many small targets at once (one per draw of a batch).
 * Every descriptor is validated before anything is launched.  The clears
 * run concurrently, so their byte spans -- per sample plane, from the first
 * to the last byte of the area -- must not overlap
 * (SWR_ERROR_INVALID_ARGUMENT otherwise): side-by-side areas of the same
 * rows count as overlapping. */
int swr_clear_batch(const swr_clear_desc *descs, uint32_t count, void *stream);

/* Compressed-texture decoding: BC_Decoder::Decode (src/Device/
 * BC_Decoder.hpp:29), which vk::Image::decodeBC calls per mip level and
 * layer (src/Vulkan/VkImage.cpp:1381-1399) to fill the decompressed image
 * the sampler reads.  `src` holds ceil(width / 4) x ceil(height / 4) blocks
 * in row-major order (8 bytes for BC1 and BC4, 16 otherwise); pixels
 * outside the image are not written.  BC1-BC3 and BC7 write whole B8G8R8A8
 * pixels (dst and its pitch 4-byte aligned, dst_bytes_per_pixel 4); BC4
 * writes byte 0 and BC5 bytes 0 and 1 of each pixel, leaving the others as
 * they are; BC6H writes four half floats R, G, B, 1.0 (dst_bytes_per_pixel
 * 8, 2-byte aligned), UF16 when no_alpha_or_unsigned is set, else SF16.
 * n is 1..7. */
typedef struct swr_bc_desc {
  const void *src;                /* device pointer, 8-byte aligned */
  void *dst;                      /* device pointer to pixel (0, 0) */
  uint32_t width, height;
  uint32_t dst_row_pitch_bytes;
  uint32_t dst_bytes_per_pixel;
  uint32_t n;                     /* the n of BCn */
  uint32_t no_alpha_or_unsigned;  /* BC1: RGB without alpha; BC4 / BC5: UNORM (else SNORM); BC6H: UF16 (else SF16) */
} swr_bc_desc;

int swr_decode_bc(const swr_bc_desc *desc, void *stream);

/* ETC2 / EAC decoding: ETC_Decoder::Decode (src/Device/ETC_Decoder.hpp:42),
 * which vk::Image::decodeETC2 calls per mip level and layer
 * (src/Vulkan/VkImage.cpp:1340-1376).  `src` holds ceil(width / 4) x
 * ceil(height / 4) blocks in row-major order (16 bytes for RG11 and RGBA8,
 * 8 otherwise); input_type is ETC_Decoder::InputType.  RGB8, RGB8 with
 * punch-through alpha and RGBA8 write B8G8R8A8 pixels (dst_bytes_per_pixel
 * 4, 4-byte aligned; a punch-through texel of index 2 is transparent
 * black); R11 writes one and RG11 two 16-bit values per pixel, the 11-bit
 * value clamped and shifted left by 5 (dst_bytes_per_pixel 2 and 4, 2-byte
 * aligned).  Pixels outside the image are not written. */
enum {
  SWR_ETC_R_SIGNED = 0,
  SWR_ETC_R_UNSIGNED = 1,
  SWR_ETC_RG_SIGNED = 2,
  SWR_ETC_RG_UNSIGNED = 3,
  SWR_ETC_RGB = 4,
  SWR_ETC_RGB_PUNCHTHROUGH_ALPHA = 5,
  SWR_ETC_RGBA = 6
};
typedef struct swr_etc_desc {
  const void *src;                /* device pointer, 8-byte aligned */
  void *dst;                      /* device pointer to pixel (0, 0) */
  uint32_t width, height;
  uint32_t dst_row_pitch_bytes;
  uint32_t dst_bytes_per_pixel;
  uint32_t input_type;            /* SWR_ETC_* */
} swr_etc_desc;

int swr_decode_etc(const swr_etc_desc *desc, void *stream);

/* Prepares `count` draws (as swr_draw_batch would render them) for
 * repeated runs; pair_headroom_percent (e.g. 10) over-provisions the pair
 * buffers beyond the planned geometry.  Synchronises `stream`. */
int swr_batch_plan_create(const swr_draw_desc *descs, uint32_t count, uint32_t pair_headroom_percent,
                          swr_batch_plan **plan, void *stream);

/* Enqueues the planned draws on `stream` (no allocation, no host wait). */
int swr_batch_plan_run(swr_batch_plan *plan, void *stream);

/* Waits for the device and reports, in *pairs_needed, the largest pair
 * count any run needed when it exceeded the plan's capacity (0: every run
 * fit).  Such runs still rendered every primitive (see swr_batch_plan). */
int swr_batch_plan_check(swr_batch_plan *plan, uint32_t *pairs_needed);

/* Times the plan under candidate tunings (tile size, raster workgroups per
 * CU), keeps the fastest and returns it in *best (may be NULL).  Renders
 * the batch several times: the attachments hold its last rendering.
 * Synchronises `stream`. */
int swr_batch_plan_autotune(swr_batch_plan *plan, swr_tuning *best, void *stream);

/* Sets the gradient targets of a plan's draws (one swr_grad_desc per
 * planned draw, in the order the draws were given; see swr_grad_desc).  The
 * buffers they name must stay allocated while the plan is used.
 * Synchronises `stream`. */
int swr_batch_plan_set_grads(swr_batch_plan *plan, const swr_grad_desc *grads, uint32_t count, void *stream);

/* Enqueues the gradients of every planned draw (as swr_draw_backward of
 * each, accumulated into the buffers swr_batch_plan_set_grads named), one
 * launch per segment and sample count: no allocation, no host wait. */
int swr_batch_plan_backward(swr_batch_plan *plan, void *stream);

/* Frees the plan (waits for the device first). */
int swr_batch_plan_destroy(swr_batch_plan *plan);

/* Gradients of the draw `desc` (see swr_grad_desc).  Asynchronous on
 * `stream`; the forward draw must have completed on the same stream. */
int swr_draw_backward(const swr_draw_desc *desc, const swr_grad_desc *grad, void *stream);

/* Fills levels 1..level_count-1 of `tex` from level 0 by 2x2 box filtering
 * (level k+1 = 0.25 * ((a + b) + (c + d)) for RGBA32F; (a+b+c+d+2)>>2 per
 * channel for RGBA8 and BGRA8; odd sizes halve with floor and clamp the second
 * texel).  The mip chain a differentiable texture is sampled through. */
int swr_generate_mips(const swr_texture *tex, void *stream);

/* Adjoint of swr_generate_mips for RGBA32F: adds each level's gradient,
 * from the coarsest up, into the level below (every texel receives 0.25 of
 * the gradient of each box it belongs to).  Level 0 then holds the gradient
 * with respect to the base texture. */
int swr_mip_backward(const swr_texture *grad, void *stream);

/* Per-stage GPU time of the draws enqueued while profiling was enabled,
 * measured with HIP events recorded on the caller's stream (the analogue
 * of the reference's timestamp queries, src/Vulkan/VkQueryPool.cpp:217). */
typedef struct swr_profile {
  double setup_ms;   /* vertex tail, clipping, setup, tile counts      */
  double bin_ms;     /* prefix sum, pair emission, sort, tile ranges   */
  double raster_ms;  /* tile rasteriser (coverage .. colour write)     */
  double total_ms;   /* first to last event of each draw, summed       */
  uint64_t draws;    /* draws measured                                 */
  uint64_t pairs;    /* (tile, triangle) pairs binned                  */
  uint64_t block_tests;    /* (8x8 block, primitive) coverage evaluations */
  uint64_t block_covers;   /* ... of which covered at least one sample    */
  uint64_t shaded;         /* fragments shaded (passed all tests)         */
} swr_profile;

/* Enables (1) or disables (0) event recording on the current device. */
int swr_profile_enable(int enable);

/* Waits for the recorded events, returns the sums since the last read
 * and resets them. */
int swr_profile_read(swr_profile *out);

/* Sets / reads the rasteriser tuning used by later draws and new batch plans
 * on this device.  The default is 16x16 tiles with one workgroup per tile. */
int swr_set_tuning(const swr_tuning *tuning);
int swr_get_tuning(swr_tuning *tuning);

#ifdef __cplusplus
}
#endif

#endif /* SWR_H_ */
