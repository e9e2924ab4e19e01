/*
 * swr.h -- C ABI of libswr, an MI355X (gfx950) sorted-wavefront path tracer.
 *
 * This is the drop-in boundary for the hot path of joshbainbridge/msc-project:
 * Pathtracer::process() (src/core/Pathtracer.cpp:425-493) and the Embree queries it makes.
 * Every entry point names the reference interface it replaces. Plain pointers and sizes only;
 * no exceptions cross the boundary. All calls return an int status (SWR_OK = 0, negative on
 * error); swr_last_error() gives the message of the last failure on a context.
 *
 * Threading: a context is single-threaded and stream-ordered; swr_cancel() is the one call
 * that may come from another thread (Pathtracer::terminate, Pathtracer.cpp:420-423).
 *
 * Pointers named *_dev are device (HBM) pointers, for example from swr_device_alloc(). All
 * other pointers are host memory, borrowed for the duration of the call only.
 */
#ifndef SWR_H
#define SWR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWR_OK 0
#define SWR_E_INVALID (-1)   /* bad argument or state */
#define SWR_E_HIP (-2)       /* HIP runtime error */
#define SWR_E_BUDGET (-3)    /* device memory budget exceeded */
#define SWR_E_CANCELLED (-4) /* swr_cancel() was called */
#define SWR_E_IO (-5)        /* file read/write failure */
#define SWR_E_PARSE (-6)     /* scene file could not be decoded */
#define SWR_E_BUSY (-7)      /* no free job slot or table entry right now; retry after jobs finish */

/* Ray masks used by the reference: camera and continuation rays see everything
 * (RayDecompress.cpp:20), shadow rays skip the lights (Integrator.cpp:156), and light
 * geometry carries 0xF0000000 (Pathtracer.cpp:267). */
#define SWR_MASK_ALL 0xFFFFFFFFu
#define SWR_MASK_SHADOW 0x0FFFFFFFu
#define SWR_MASK_LIGHT 0xF0000000u

#define SWR_CAMERA_PINHOLE 0   /* PinHoleCamera.cpp */
#define SWR_CAMERA_THINLENS 1  /* ThinLensCamera.cpp */
#define SWR_SAMPLER_STRATIFIED 0 /* StratifiedSampler.cpp */
#define SWR_SAMPLER_INDEPENDENT 1 /* IndependentSampler.cpp */
#define SWR_SAMPLER_GRID 2     /* GridSampler.cpp */
#define SWR_FILTER_TENT 0      /* TentFilter.cpp */
#define SWR_FILTER_BOX 1       /* BoxFilter.cpp */
#define SWR_SHADER_LAMBERT 0   /* LambertShader.cpp */
#define SWR_SHADER_NULL 1      /* NullShader.cpp */
#define SWR_SHADER_GLOSSY 2    /* GGX reflection lobe (no reference counterpart) */

typedef struct swr_context swr_context;
typedef struct swr_scene swr_scene;

/* Camera block of the setup document (PinHoleCamera.h / ThinLensCamera.h convert<>). */
typedef struct {
  int32_t type;
  float translation[3];
  float rotation[3];     /* degrees, camera order T*Ry*Rx*Rz (PinHoleCamera.cpp:7-11) */
  float focal_length;    /* mm */
  float focal_distance;  /* ThinLens only */
  float aperture;        /* f-number, ThinLens only */
} swr_camera_desc;

/* Image, settings, sampler, filter and camera of the setup document (Image.h:37-50,
 * Settings.h:22-37, Pathtracer.cpp:45-161). */
typedef struct {
  int32_t width, height;
  int32_t base;          /* sample base: spp = base*base (Image.h:46, Camera.cpp:10) */
  int32_t spp;           /* 0 -> base*base; otherwise explicit spp (Independent sampler) */
  int32_t sampler;
  int32_t filter;
  int32_t min_depth, max_depth;
  float threshold;
  int32_t shading_size;  /* accepted for schema parity; shading ranges come from the hit sort */
  int32_t bin_exponent;  /* wave capacity = 2^bin_exponent rays (DirectionalBins.cpp:15) */
  uint32_t iteration;    /* progressive iteration (Image.h:27): an RNG counter word */
  uint64_t seed;         /* Philox-4x32-10 key for this render */
  swr_camera_desc camera;
} swr_render_desc;

/* Shader records (LambertShader.h / NullShader.h convert<>; ConstantTexture colour). */
typedef struct {
  int32_t type;
  float colour[3];
  float reflectance;
  float roughness;       /* Glossy only: GGX alpha */
} swr_shader_desc;

/* Texture nodes: the reference's TextureInterface implementations. CONSTANT is ConstantTexture
 * (ConstantTexture.h convert<>); IMAGE is StandardTexture (StandardTexture.cpp:11-43), an image
 * looked up at the hit's interpolated texture coordinates with periodic wrap in s and t and zero
 * filter width, as OpenImageIO's texture() does there: texel centres at (i + 0.5) / size and
 * uniform cubic B-spline weights over the 4 x 4 texels around (s, t); LAYERED is LayeredTexture
 * (LayeredTexture.cpp:14-32): upper * mask + lower * (1 - mask). At most 8 colours may be
 * pending while one shader's texture graph is evaluated (each LAYERED node holds its upper and
 * lower colours while its mask is evaluated); deeper graphs fail at commit. */
#define SWR_TEXTURE_CONSTANT 0
#define SWR_TEXTURE_IMAGE 1
#define SWR_TEXTURE_LAYERED 2
typedef struct {
  int32_t type;
  float colour[3];              /* CONSTANT */
  int32_t image;                /* IMAGE: index returned by swr_scene_add_image */
  int32_t upper, lower, mask;   /* LAYERED: indices of earlier texture nodes */
} swr_texture_desc;

/* Quad light record (QuadLight.h convert<>). */
typedef struct {
  float translation[3];
  float rotation[3];     /* degrees, light order T*Rx*Ry*Rz*S (QuadLight.cpp:24-29) */
  float scale[2];
  float intensity;
} swr_quad_light_desc;

/* Per-render counters and stage times. */
typedef struct {
  uint64_t camera_rays;
  uint64_t closest_rays;   /* closest-hit queries traced (camera + continuation) */
  uint64_t shadow_rays;    /* occlusion queries traced */
  uint32_t waves;          /* wavefront iterations */
  uint32_t max_wave;       /* largest wave (rays) */
  double ms_total;         /* first raygen to resolved framebuffer, device time */
  /* Device time per stage. Radiance is added to each sample in place by the shading and shadow
   * kernels (Integrator.cpp:66-71, 171-173), so accumulation is part of ms_shade / ms_shadow. */
  double ms_raygen, ms_sort_rays, ms_trace, ms_sort_hits, ms_shade, ms_shadow, ms_filter;
} swr_stats;

/* Pass structures of the radix sort: onesweep (one kernel per digit, decoupled look-back over
 * earlier tiles), reduce-then-scan (per digit: tile counts, a scan over tiles, the scatter;
 * no tile waits for another), or the bucket sort (one stable partition of the key range into
 * 2048 buckets, then each bucket sorted on its remaining bits in LDS: two trips through device
 * memory instead of one per 8-bit digit; used for sorts of more than 16 key bits whose count is
 * known at launch, reduce-then-scan otherwise). */
#define SWR_SORT_ONESWEEP 0
#define SWR_SORT_REDUCE_SCAN 1
#define SWR_SORT_BUCKET 2

/* Launch shapes of the loop. The reference exposes the equivalent knobs as settings
 * (bucket size, shading size, bin exponent; Settings.h:16-24); these are chosen per context by
 * swr_autotune or set directly. */
typedef struct {
  int32_t sort_keys_per_thread;  /* sort tile = 512 * this: 8, 12, 16, 20 or 24 */
  int32_t ray_key_bits;          /* top bits of the octant + Morton ray key the ray sort orders by (3..32) */
  int32_t hit_key_bits;          /* top bits of the hit key the hit sort orders by (0 = all bits) */
  int32_t trace_stack_entries;   /* LDS ring of the traversal stack per ray: 8, 16, 32 or 64 entries, deeper
                                    entries spill to device memory (0 = per scene: the ring
                                    swr_batch_autotune chose for it, else sized from its BVH depth,
                                    at most 16) */
  int32_t sort_algorithm;        /* SWR_SORT_ONESWEEP, SWR_SORT_REDUCE_SCAN or SWR_SORT_BUCKET */
  int32_t shadow_overlap;        /* 1 (swr_create's default): each wave's shadow pass runs on a second
                                    stream beside the next wave's traversal; 0: in line after its
                                    shading. The same images either way. */
} swr_tuning;

/* ---- context (Pathtracer ctor/dtor, Pathtracer.cpp:369-392) ------------------------------ */
int swr_create(int device, swr_context** out);
/* Frees everything the context owns. Scenes and batches still alive lose their device memory and
 * their context: their handles stay valid for swr_scene_destroy / swr_batch_destroy, every other
 * call on them fails with SWR_E_INVALID. Destroy servers before their context. */
void swr_destroy(swr_context* ctx);
const char* swr_last_error(const swr_context* ctx);
/* Device-memory budget for this context in bytes (0 = unlimited). Allocations beyond it fail
 * with SWR_E_BUDGET. */
int swr_set_budget(swr_context* ctx, uint64_t bytes);
uint64_t swr_bytes_in_use(const swr_context* ctx);
/* Thread-safe cooperative cancel (Pathtracer::terminate, Pathtracer.cpp:420-423): the render in
 * flight returns SWR_E_CANCELLED at its next wave boundary. swr_reset_cancel() re-arms. */
int swr_cancel(swr_context* ctx);
int swr_reset_cancel(swr_context* ctx);
int swr_synchronize(swr_context* ctx);
/* Device milliseconds of the last primitive entry point (sort, keys, gather, intersect,
 * occluded, raygen), measured with HIP events on the context stream around its kernels. */
double swr_last_device_ms(const swr_context* ctx);
int swr_get_tuning(const swr_context* ctx, swr_tuning* out);
int swr_set_tuning(swr_context* ctx, const swr_tuning* tuning);
/* Times candidate tunings on a reduced render of desc (same scene, same camera, about
 * 1/16 of the samples) and keeps the fastest on the context; *out receives it. */
int swr_autotune(swr_context* ctx, const swr_scene* scene, const swr_render_desc* desc, swr_tuning* out);

/* ---- device memory helpers (plumbing for callers without their own allocator) ------------- */
int swr_device_alloc(swr_context* ctx, size_t bytes, void** out_dev);
int swr_device_free(swr_context* ctx, void* dev);
int swr_memcpy_h2d(swr_context* ctx, void* dst_dev, const void* src, size_t bytes);
int swr_memcpy_d2h(swr_context* ctx, void* dst, const void* src_dev, size_t bytes);

/* ---- sort and key primitives (stages of Pathtracer::process) ------------------------------ */
/* Stable radix sort of (key, value) pairs in place, as the context tuning's sort_algorithm says
 * (onesweep on a new context; swr_autotune may choose another): LSD passes of 8-bit digits, each
 * onesweep with decoupled look-back or reduce-then-scan, or the bucket sort (a partition of the
 * key range into 2048 buckets, then every bucket sorted in LDS) for keys of more than 16 bits.
 * Replaces the ray sort (RaySort.cpp:7-118, Pathtracer.cpp:321-327) and the hit sort
 * (tbb::parallel_sort with CompareHit, Pathtracer.cpp:335-339). key_bits (1..32) is the key
 * width: keys must lie below 2^key_bits, and the sort orders them by their low key_bits bits.
 * With iota_values = 1 the input values are not read and the output
 * values are the stable permutation (original index of each sorted key). n <= 2^30. */
int swr_sort_pairs_u32(swr_context* ctx, uint32_t* keys_dev, uint32_t* vals_dev, size_t n, int key_bits,
                       int iota_values);
/* The same sort (same arguments, same result) without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its workspace is a file-scope device array of the library
 * (about 8.3 MiB per device, held until the process ends): at most 2^20 pairs per call, larger n
 * is refused with SWR_E_INVALID. Always the reduce-then-scan pass structure, at the context's
 * keys-per-lane tuning. Every call on one device shares that array, so calls that overlap on a
 * device (two contexts, two threads) race; calls on one context are stream-ordered and safe. */
int swr_sort_pairs_u32_noalloc(swr_context* ctx, uint32_t* keys_dev, uint32_t* vals_dev, size_t n, int key_bits,
                               int iota_values);
/* Ray binning key: octant(3 bits) << 29 | morton29(quantised origin), the octant+Morton
 * replacement for the cardinal bins (Camera.cpp:42-46, Integrator.cpp:240-244).
 * rays_dev: n x 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax). scale[i] = 2^b_i/(hi_i-lo_i) with
 * b = (10,10,9). */
int swr_ray_keys(swr_context* ctx, const float* rays_dev, size_t n, const float lo[3], const float scale[3],
                 uint32_t* keys_dev);
/* Hit shading key = global triangle id + 1, misses 0: the (geomID, primID) order of CompareHit
 * (RayUncompressed.h:54-60) with misses first. hits_dev: n x 4 words (t,u,v,prim). */
int swr_hit_keys(swr_context* ctx, const uint32_t* hits_dev, size_t n, uint32_t* keys_dev);
/* Synthetic hit-record key ((mesh+1) << 21) | face for mesh < 2^10, face < 2^21. */
int swr_mesh_face_keys(swr_context* ctx, const uint32_t* mesh_dev, const uint32_t* face_dev, size_t n,
                       uint32_t* keys_dev);
/* Gather: out[i] = in[perm[i]] for records of rec_words 32-bit words (1..16). */
int swr_gather_u32(swr_context* ctx, const uint32_t* in_dev, const uint32_t* perm_dev, size_t n, int rec_words,
                   uint32_t* out_dev);

/* ---- scene (Pathtracer::construct scene part, Pathtracer.cpp:163-274) --------------------- */
int swr_scene_create(swr_context* ctx, swr_scene** out);
void swr_scene_destroy(swr_scene* scene);
/* Shaders in document order; returns the shader index used by swr_scene_add_mesh. */
int swr_scene_add_shader(swr_scene* scene, const swr_shader_desc* desc);
/* An RGB float image for IMAGE textures (the OpenImageIO texture file of StandardTexture, decoded):
 * width x height texels, rows from t = 0 down, 3 floats each. Returns the image index. */
int swr_scene_add_image(swr_scene* scene, int32_t width, int32_t height, const float* rgb);
/* A texture node (LAYERED nodes name earlier nodes); returns the node index. */
int swr_scene_add_texture(swr_scene* scene, const swr_texture_desc* desc);
/* Gives a shader a texture (LambertShader's texture, Integrator.cpp:80-99): its colour at a hit is
 * the texture at (u, v) = (1 - s - t) uv0 + s uv1 + t uv2 of the hit triangle's vertex uvs and hit
 * barycentrics (PolygonObject.cpp:45-63), instead of the shader's constant colour. */
int swr_scene_set_shader_texture(swr_scene* scene, int32_t shader, int32_t texture);
/* rtcNewTriangleMesh + rtcSetBuffer (Pathtracer.cpp:178-201): float4 positions (w ignored),
 * u32x3 indices, optional float2 uv per vertex. Returns the object index (= geomID). */
int swr_scene_add_mesh(swr_scene* scene, const float* pos4, uint32_t nverts, const uint32_t* idx3, uint32_t ntris,
                       const float* uv2, int32_t shader);
/* Quad light (Pathtracer.cpp:219-272): a placeholder object with a Null shader and mask
 * 0xF0000000 is appended after all objects at commit, as in the reference. */
int swr_scene_add_quad_light(swr_scene* scene, const swr_quad_light_desc* desc);
/* Constant environment radiance seen by rays that miss (the reference's misses add nothing,
 * Integrator.cpp:26-27; default 0). */
int swr_scene_set_environment(swr_scene* scene, const float rgb[3]);
/* rtcCommit (Pathtracer.cpp:274): uploads the geometry and builds the BVH4 on device (Morton
 * radix tree, treelet restructuring, SAH-optimal collapse; about 125 ms for 2.1 M triangles). */
int swr_scene_commit(swr_scene* scene);
/* Decodes a reference two-document YAML scene file (Pathtracer.cpp:32-275) into a committed
 * scene plus its render description. OBJ paths resolve relative to the YAML file. */
int swr_scene_load_yaml(swr_context* ctx, const char* path, swr_scene** out_scene, swr_render_desc* out_desc);
uint32_t swr_scene_num_triangles(const swr_scene* scene);
uint32_t swr_scene_num_geometries(const swr_scene* scene);
/* BVH facts for tests: nodes, leaves, max depth, and a validity check (every triangle in
 * exactly one leaf; child boxes inside parent boxes) -> *valid = 1. */
int swr_scene_bvh_info(const swr_scene* scene, uint32_t* nodes, uint32_t* max_depth, uint32_t* valid);
/* Scene bounds (lo xyz, hi xyz). */
int swr_scene_bounds(const swr_scene* scene, float out6[6]);

/* ---- traversal (Embree queries) ------------------------------------------------------------ */
/* rtcIntersect (RayIntersect.cpp:7-12): closest hit in (tmin, tmax) honouring ray & geometry
 * masks. hits_dev: n x 4 words (t, u, v, prim); prim = global triangle id (geomID-major,
 * creation order) or 0xFFFFFFFF; on a miss t = tmax. */
int swr_intersect(swr_context* ctx, const swr_scene* scene, const float* rays_dev, size_t n, uint32_t mask,
                  uint32_t* hits_dev);
/* rtcOccluded (Integrator.cpp:153-174): flags_dev[i] = 1 if any hit in (tmin, tmax). */
int swr_occluded(swr_context* ctx, const swr_scene* scene, const float* rays_dev, size_t n, uint32_t mask,
                 uint32_t* flags_dev);
/* Traversal stack ring of this scene (0 = sized from its BVH depth, at most 16); swr_batch_autotune
 * sets it.
 * The context's trace_stack_entries, when nonzero, overrides it. Rings: 0, 8, 16, 32 or 64. */
int swr_scene_set_stack_ring(swr_scene* scene, int32_t ring);
int swr_scene_stack_ring(const swr_scene* scene, int32_t* ring);
/* Global triangle id -> (geomID, primID). */
int swr_scene_prim_to_geom(const swr_scene* scene, uint32_t prim, uint32_t* geom_id, uint32_t* prim_id);

/* ---- render (Pathtracer::process, Pathtracer.cpp:425-493) ---------------------------------- */
/* Camera rays for sample ids [sid0, sid0+n) (Camera.cpp:6-59) into rays_dev (n x 8 floats) and
 * film positions into xy_dev (n x 2 floats, may be NULL). */
int swr_raygen(swr_context* ctx, const swr_render_desc* desc, uint32_t sid0, size_t n, float* rays_dev,
               float* xy_dev);
/* One iteration: every sample traced through the sorted wavefront loop, reduced per sample and
 * filtered. pixels: host W*H*3 floats, the filtered iteration (row-major y*W+x as Pixel in
 * Image.h:16-20), or NULL to leave it in device memory; samples: optional host W*H*spp*5 floats
 * (x, y, r, g, b) in reference sampleID order (Camera.cpp:28). stats may be NULL.
 * Out of core (the role of the reference's bin files, DirectionalBins.cpp:37-115): when the
 * frame's per-sample state (24 B per sample, 44 B with samples requested) does not fit the
 * context budget, the frame is traced as column bands one after another, each band holding its
 * columns plus the tent filter's neighbour columns; the image and sample records are identical
 * to an unbanded render. SWR_E_BUDGET if the framebuffer or one column of samples cannot fit. */
int swr_render(swr_context* ctx, const swr_scene* scene, const swr_render_desc* desc, float* pixels, float* samples,
               swr_stats* stats);

/* ---- batched renders (Pathtracer::process of many cameras in one loop) ----------------------
 * The reference renders one image per Pathtracer and drains its bins batch by batch
 * (Pathtracer.cpp:451-475). A batch traces the rays of many render jobs (any resolution, spp
 * and resident scene) through one persistent wavefront loop: every wave is filled from all jobs,
 * the per-wave kernels are captured once as HIP graphs and replayed without host round trips.
 * Each job's image is bit-identical to swr_render of the same job alone. */
typedef struct swr_batch swr_batch;
int swr_batch_create(swr_context* ctx, swr_batch** out);
void swr_batch_destroy(swr_batch* batch);
/* Adds a job (at most 65536; scenes: at most 256 distinct). *job receives its index. */
int swr_batch_add(swr_batch* batch, const swr_scene* scene, const swr_render_desc* desc, int32_t* job);
/* Traces every job to completion with waves of wave_rays rays (0 = 2^23, at most 2^28). stats: the whole
 * batch (ms_total = first raygen to the last resolved framebuffer). Out of core: when the
 * context budget (swr_set_budget) cannot hold every job's state (24 B per sample, 12 B per
 * pixel), the jobs run in groups of consecutive jobs that fit, one loop per group, each group's
 * images copied to host memory as it finishes (same images; sample records are then not kept);
 * SWR_E_BUDGET if one job cannot fit alone. flags: 0, or
 * SWR_BATCH_PACED to launch each wave from the host with sizes read back after the previous one
 * (no graphs; fills the per-stage times; for profiling and comparison). */
#define SWR_BATCH_PACED 1u
int swr_batch_run(swr_batch* batch, uint32_t wave_rays, uint32_t flags, swr_stats* stats);
/* A finished job's filtered pixels (W*H*3 floats) and optionally its samples (as swr_render). */
int swr_batch_result(swr_batch* batch, int32_t job, float* pixels, float* samples);
/* Chooses the traversal stack ring per scene (recorded on each scene, used by every later loop
 * while the context's trace_stack_entries is 0), then the sort tile and key widths, on reduced
 * copies of the batch's jobs (about 1/16 of their samples). */
int swr_batch_autotune(swr_batch* batch, uint32_t wave_rays, swr_tuning* out);

/* ---- serving loop (progressive renders on resident scenes, continuous batching) ------------
 * The reference's viewer drives one Pathtracer, calling process() again and again and showing
 * pixels / iteration (main.cpp:104-124, 158-173). A server keeps scenes resident and runs one
 * persistent wavefront loop on its own thread; requests (a camera on a scene, M iterations)
 * are admitted as job slots and device memory free up, and their waves merge with everything
 * else in flight. Admission keeps a short backlog of ungenerated paths and fills it round
 * robin, one iteration per admitted request per turn (a new request's first iteration at once),
 * so short requests pass long ones and long ones are never starved. A request's result is the mean of its M filtered iterations, accumulated in
 * iteration order -- bit-identical to the same request served alone.
 * While a server exists its thread owns the context: use only swr_server_* and swr_cancel. */
typedef struct swr_server swr_server;
typedef struct {
  uint64_t state_bytes;      /* device memory for request state (per-sample arenas, 24 B per sample of one
                                iteration, and FP32 framebuffers, 12 B per pixel); 0 = 8 GiB */
  uint32_t wave_rays;        /* rays per wave (0 = 2^23, at most 2^26) */
  uint32_t max_requests;     /* job slots: iterations on the device at once (0 = 256, at most 65536) */
  uint32_t waves_per_check;  /* graph-replayed waves between host checks (0 = 2) */
  uint32_t iteration_window; /* iterations of one request in the loop at once (0 = 16); each holds its own
                                per-sample arena range */
  uint32_t backlog_waves;    /* admit while fewer paths than this many waves wait to be generated (0 = 2) */
} swr_server_config;
typedef struct {
  int32_t status;            /* SWR_OK, or why the request failed */
  uint32_t iterations;
  double queue_ms;           /* submission to admission */
  double latency_ms;         /* submission to the framebuffer in host memory */
} swr_request_stats;
typedef struct {
  uint64_t submitted, completed, refused;
  uint64_t camera_rays, closest_rays, shadow_rays, waves;
  double busy_s;             /* wall time with at least one request on the device */
  uint64_t state_bytes, queue_bytes;  /* request-state arenas; ray queues and tables */
} swr_server_totals;
/* scenes: the resident scenes (their order gives the scene index of a request). */
int swr_server_create(swr_context* ctx, const swr_scene* const* scenes, int32_t num_scenes,
                      const swr_server_config* config, swr_server** out);
/* Queues a request: desc gives one iteration (image size, spp, sampler, filter, camera, depth,
 * seed; iteration i uses desc->iteration + i as its RNG counter word) and iterations its count.
 * pixels (host, W*H*3 floats) receives the result and must stay valid until swr_server_wait
 * returns; it may be NULL. Thread-safe. An iteration whose per-sample state exceeds the server's
 * per-sample arena runs as column bands (each a job of its own, at most half the arena each,
 * the tent's neighbour columns traced by both neighbours), with the same result as unbanded.
 * SWR_E_BUDGET if the framebuffer or one column of samples can never fit. */
int swr_server_submit(swr_server* server, int32_t scene, const swr_render_desc* desc, int32_t iterations,
                      float* pixels, uint64_t* request);
/* Blocks until the request has finished (or failed); thread-safe. */
int swr_server_wait(swr_server* server, uint64_t request, swr_request_stats* stats);
int swr_server_stats(swr_server* server, swr_server_totals* stats);
/* Stops the thread; requests not finished fail with SWR_E_CANCELLED. */
void swr_server_destroy(swr_server* server);

/* ---- file output (main.cpp:197-208) ------------------------------------------------------ */
/* Writes W*H*3 floats as an uncompressed scanline OpenEXR file of half RGBA with A = 1. */
int swr_write_exr(const char* path, const float* rgb, int32_t width, int32_t height);

#ifdef __cplusplus
}
#endif
#endif
