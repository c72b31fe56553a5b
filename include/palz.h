/*
 * palz.h — C ABI of the MI355X (gfx950) caustic hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (PALZ_OK == 0) and is stream ordered on the `stream` argument
 * (a hipStream_t passed as void*; NULL = the legacy default stream).  Device
 * pointers are caller-owned unless stated otherwise.  Handles are not
 * thread-safe; use one handle per host thread.  palz_last_error() returns a
 * thread-local message for the last non-zero status.
 *
 * The POD records below reproduce the reference's in-memory layouts byte for
 * byte so that buffers can be exchanged with code that uses the reference
 * structs (sizes are static_asserted in the implementation).
 */
#ifndef PALZ_H
#define PALZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    PALZ_OK = 0,
    PALZ_ERR_INVALID_ARGUMENT = 1,
    PALZ_ERR_HIP = 2,            /* a HIP runtime call failed                    */
    PALZ_ERR_OUT_OF_MEMORY = 3,
    PALZ_ERR_CAPACITY = 4,       /* an output or scratch capacity was exceeded   */
    PALZ_ERR_TRAVERSAL_STACK = 5 /* a BVH traversal needed a deeper stack        */
};

/* ---- POD records (layouts of the reference) ---------------------------- */

/* float3: 12 B, no padding (helper_math.h float3). */
typedef struct { float x, y, z; } palz_float3;
typedef struct { float x, y, z, w; } palz_float4;

/* minimum_lbvh::Triangle (minimum_lbvh.h:207-210): 36 B. */
typedef struct { palz_float3 vs[3]; } palz_triangle;

/* minimum_lbvh::AABB (minimum_lbvh.h:220-261): 24 B. */
typedef struct { palz_float3 lower, upper; } palz_aabb;

/* minimum_lbvh::NodeIndex (minimum_lbvh.h:263-274): bits [0,31) index, bit 31
 * leaf flag.  The invalid node is 0x7FFFFFFF (index 0x7FFFFFFF, not a leaf). */
typedef uint32_t palz_node_index;
#define PALZ_NODE_INVALID 0x7FFFFFFFu
#define PALZ_NODE_LEAF_BIT 0x80000000u

/* minimum_lbvh::InternalNode (minimum_lbvh.h:285-291): 64 B.
 * `context` holds the escape link after a build: the upper bound of the
 * node's leaf range when that is < n-1, else 0xFFFFFFFF. */
typedef struct {
    palz_node_index children[2];
    palz_aabb aabbs[2];
    palz_node_index parent;
    uint32_t context;
} palz_internal_node;

/* A two-level (BVH4) view of internal node i of an LBVH, for packet
 * traversal: its two children (as in the node) and, for each internal child
 * c, that child's two children at grand[2c + j] (PALZ_NODE_INVALID below a
 * leaf child), with the boxes of all six by plane: plane[a][0][s] is box s's
 * lower bound on axis a, plane[a][1][s] its upper bound and plane[a][2][s]
 * the lower bound again, for s = 0, 1 the children's boxes and s = 2 + 2c + j
 * the grandchildren's (those below a leaf child hold 0).  So plane[a][0..1]
 * and plane[a][1..2] are both (entry, exit) pairs: for rays whose direction
 * component a is > 0 and < 0, a wave loads one or the other.
 * No counterpart in the reference: it holds what intersect_escape_link
 * (minimum_lbvh.h:1078-1152) reads from three nodes, so that one load serves
 * two levels.  248 B. */
typedef struct {
    palz_node_index child[2];
    palz_node_index grand[4];
    uint32_t pad_[2];
    float plane[3][3][6];
} palz_wide_node;

/* Material (pk.h:67-72) and TriangleAttrib (pk.h:74-78): 40 B. */
enum { PALZ_MATERIAL_DIFFUSE = 0, PALZ_MATERIAL_MIRROR = 1, PALZ_MATERIAL_DIELECTRIC = 2 };
typedef struct {
    int32_t material;
    palz_float3 shading_normals[3];
} palz_triangle_attrib;

/* BUILD_OPTION (minimum_lbvh.h:42-46). USE_NORMAL follows the GPU packing of
 * minimum_lbvh.cu:85; CPU_PARALLEL has no meaning on the GPU and is ignored. */
enum { PALZ_BUILD_OPTION_USE_NORMAL = 1, PALZ_BUILD_OPTION_CPU_PARALLEL = 2 };

/* ---- library ------------------------------------------------------------ */
const char* palz_last_error(void);
int palz_version(void);

/* ---- device memory, streams, events ---------------------------------------
 * Thin wrappers over the HIP runtime so that hosts without a HIP binding
 * (ctypes, cgo, JNI) can drive the library.  They stand in for the
 * reference's TypedBuffer / oroMalloc / DeviceStopwatch plumbing
 * (typedbuffer.h:19-146, pk.h:1192-1223). */
int palz_device_count(int* count);
int palz_set_device(int device);
int palz_malloc(void** d_ptr, size_t bytes);
int palz_free(void* d_ptr);
int palz_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream);
int palz_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream);
int palz_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream);
int palz_memset(void* d_dst, int value, size_t bytes, void* stream);
int palz_stream_create(void** stream);
int palz_stream_destroy(void* stream);
int palz_stream_sync(void* stream);
int palz_device_sync(void);
int palz_event_create(void** event);
int palz_event_destroy(void* event);
int palz_event_record(void* event, void* stream);
int palz_event_elapsed_ms(float* ms, void* start, void* stop);
/* A plain streaming copy kernel (16-byte aligned, size a multiple of 16): the
 * yardstick for achievable HBM bandwidth and for HBM counter calibration. */
int palz_copy_bytes(void* d_dst, const void* d_src, size_t bytes, void* stream);

/* ---- onesweep radix sort -------------------------------------------------
 * Replaces tinyhiponesweep::OnesweepSort (tinyhiponesweep.h:572-668).
 * The sorter owns its scratch (digit histograms, look-back table, tile
 * tickets) and grows it on demand. */
typedef struct palz_sorter palz_sorter;
int palz_sorter_create(palz_sorter** out);
int palz_sorter_destroy(palz_sorter* s);
/* Pre-size the look-back table for n keys.  Sorting more keys than reserved
 * grows it with hipMalloc, so reserve before capturing a sort in a hipGraph.
 * (The reference sizes its table once, tinyhiponesweep.h:725-732.) */
int palz_sorter_reserve(palz_sorter* s, uint32_t n);

/* Stable LSD sort of 64-bit words on bit range [start_bit, end_bit), 8-bit
 * digits.  Replaces OnesweepSort::sort(const KeyValueSoA64&, ..., n, startBit,
 * endBit, stream) with `value == nullptr` (tinyhiponesweep.h:596-599), which is
 * how both reference call sites use it (minimum_lbvh.h:848, pk.h:1526).  The
 * result is written back to `words`; `tmp` (n words) is clobbered.  If
 * `values` is not NULL the u32 payload moves with its word (`tmp_values`
 * must then hold n entries).  n must be < 2^30. */
int palz_sort_u64(palz_sorter* s, uint64_t* words, uint64_t* tmp, uint32_t* values,
                  uint32_t* tmp_values, uint32_t n, int start_bit, int end_bit, void* stream);

/* Stable LSD sort of 32-bit keys on bit range [start_bit, end_bit) (end_bit
 * <= 32), 8-bit digits, with an optional u32 payload.  Replaces
 * OnesweepSort::sort(const KeyValueSoA32&, const KeyValueSoA32&, n, startBit,
 * endBit, stream) (tinyhiponesweep.h:562-566, :592-595): `keys`/`values` are
 * KeyValueSoA32's key/value arrays, `tmp`/`tmp_values` its tmp's.  The result
 * is written back to `keys` (and `values`).  n must be < 2^30. */
int palz_sort_u32(palz_sorter* s, uint32_t* keys, uint32_t* tmp, uint32_t* values, uint32_t* tmp_values,
                  uint32_t n, int start_bit, int end_bit, void* stream);

/* palz_sort_u64 without a sorter and without allocating: no counterpart in
 * the reference.  For callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as palz_sort_u64; the scratch (look-back
 * table, digit histograms, tile tickets) is a file-scope device array of
 * 8.4 MB that each device holds until the process ends.  n must be <= 2^24
 * (PALZ_ERR_CAPACITY otherwise).  The first call on a device sets the array
 * up and synchronises `stream`.  Every call on a device shares its array:
 * calls that overlap on one device (other streams or threads) race. */
int palz_sort_u64_noalloc(uint64_t* words, uint64_t* tmp, uint32_t* values, uint32_t* tmp_values, uint32_t n,
                          int start_bit, int end_bit, void* stream);

/* ---- LBVH builder ---------------------------------------------------------
 * Replaces minimum_lbvh::BVHGPUBuilder (minimum_lbvh.h:741-900).  The builder
 * owns the node array (n-1 records) and the root index, as the reference does
 * (m_internals, m_rootNode). */
typedef struct palz_lbvh palz_lbvh;
int palz_lbvh_create(palz_lbvh** out);
int palz_lbvh_destroy(palz_lbvh* b);

/* BVHGPUBuilder::build(triangles, nTriangles, buildOption, sorter, stream)
 * (minimum_lbvh.h:798).  morton_bits = 32 gives the reference key (top 32 bits
 * of a 63-bit Morton code, minimum_lbvh.cu:89); 30 gives a 10-bit-per-axis
 * Morton key.  d_tris is a device array of n triangles. */
int palz_lbvh_build(palz_lbvh* b, const palz_triangle* d_tris, int n, uint32_t build_option,
                    int morton_bits, palz_sorter* sorter, void* stream);

/* Device pointers to the result (valid until the next build or destroy).
 * *d_nodes has n-1 records, *d_root one palz_node_index, *d_scene_aabb one
 * palz_aabb.  Any out-pointer may be NULL. */
int palz_lbvh_result(palz_lbvh* b, palz_internal_node** d_nodes, palz_node_index** d_root,
                     palz_aabb** d_scene_aabb, int* n_triangles);

/* Device pointer to the sorted {key, index} words of the last build (n words,
 * key in bits [0,32), triangle index in bits [32,64)) and the n-1 deltas;
 * exposed for parity checks. */
int palz_lbvh_sorted(palz_lbvh* b, uint64_t** d_words, uint8_t** d_deltas);

/* The BVH4 view of the last build (n-1 records, one per internal node;
 * undefined for n = 1), computed on `stream` by this call and owned by the
 * builder like the nodes (valid until the next build or destroy).  No
 * counterpart in the reference; the batched solver's occlusion rays use it
 * when a scene descriptor names it. */
int palz_lbvh_widen(palz_lbvh* b, palz_wide_node** d_wide, void* stream);

/* InternalNormalBound (pk.h:1705-1709): per internal node, the AABB of the
 * shading normals of every triangle under each child, and the arrival
 * counter of the bottom-up pass (2 for every node afterwards).  52 B. */
typedef struct {
    palz_aabb normal_bounds[2];
    uint32_t counter;
} palz_normal_bound;

/* The bottom-up normal-bound propagation of main.cpp:248-295 over the last
 * build's tree: d_attribs[t] is triangle t's TriangleAttrib; d_out receives
 * n-1 records (nothing for n = 1).  Boxes take the canonical min/max, so they
 * do not depend on the order the subtrees arrive in. */
int palz_lbvh_normal_bounds(palz_lbvh* b, const palz_triangle_attrib* d_attribs, palz_normal_bound* d_out,
                            void* stream);

/* Stage timing (replaces the commented-out DeviceStopwatch of
 * minimum_lbvh.h:800-801, :884-885).  When enabled, each build records HIP
 * events on its stream after every stage; palz_lbvh_timings waits for the
 * last one and returns up to 8 durations in ms: [0] scene AABB, [1] Morton
 * keys + digit counts, [2..5] the four sort passes, [6] deltas and the leaf
 * level of the build, [7] the levels above it. */
int palz_lbvh_set_timing(palz_lbvh* b, int enable);
int palz_lbvh_timings(palz_lbvh* b, float* ms, int max_stages, int* n_stages);

/* ---- scenes, photons, tuples, paths ------------------------------------ */

/* FirstDiffuse (pk.h:80-86): a receiver point.  R == (0,0,0) marks it
 * invalid (main_gpu.cu:385); lambda is ignored by the solvers here, which
 * take an explicit wavelength list. */
typedef struct {
    palz_float3 p;
    palz_float3 ng;
    palz_float3 R;
    float lambda;
} palz_first_diffuse;

/* Device pointers of one scene (triangles, attributes and its BVH). */
typedef struct {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    const palz_node_index* root;
    uint32_t n_tris;
    uint32_t pad_;
} palz_scene_view;

/* minimum_lbvh::Hit (minimum_lbvh.h:905-911): t (the query's t_max when
 * nothing is hit), barycentrics (u, v) of vertices 1 and 2, the unnormalised
 * geometric normal cross(v1 - v0, v2 - v1), triangle index (0xFFFFFFFF when
 * nothing is hit). */
typedef struct {
    float t;
    float u, v;
    palz_float3 ng;
    uint32_t tri;
} palz_hit;

enum {
    PALZ_QUERY_CLOSEST = 0,           /* intersect_stackfull (minimum_lbvh.h:927-988) */
    PALZ_QUERY_CLOSEST_STACKFREE = 1, /* intersect_stackfree, RAY_QUERY_CLOSEST (:996-1076) */
    PALZ_QUERY_ANY = 2                /* intersect_stackfree, RAY_QUERY_ANY: first hit found */
};

/* Casts n rays (d_ro, d_rd: 3 floats per ray, device memory) into the scene's
 * BVH over [0, t_max] and writes one palz_hit per ray.  Replaces the
 * reference's traversal device functions (minimum_lbvh.h:927-1076) for a
 * caller that holds rays rather than a kernel of its own.  Closest-hit
 * queries take t_max = FLT_MAX as the reference does. */
int palz_intersect_rays(const palz_scene_view* view, const float* d_ro, const float* d_rd, uint32_t n, int query,
                        float t_max, palz_hit* d_hits, void* stream);

/* photonTrace arguments (main_gpu.cu:56, pk.h:1441-1459). */
typedef struct {
    palz_float3 p_light;
    float eta_min;               /* n(390 nm) in the reference: lerp end points */
    float eta_max;
    int32_t iteration;
    float min_throughput;        /* m_minThroughput = 0.05 (pk.h:1687) */
    uint32_t photons_per_triangle;  /* the reference launches 256 (pk.h:1456-1457) */
} palz_photon_params;

/* One photon's specular tuple: PathCache::TrianglePath (pk.h:234-240) plus the
 * photon id and landing point.  Records are dense by photon id =
 * triangle * photons_per_triangle + j; event == 0 marks "no tuple". */
typedef struct {
    uint32_t photon;
    uint32_t event;       /* EventDescriptor::asU32 (pk.h:156-158) */
    int32_t tris[4];
    float params[8];
    palz_float3 p_final;
    float throughput;
} palz_photon_record;

/* One (receiver point, tuple) work item: SpecularPath (pk.h:1102-1106) with
 * the cached tuple and its initial barycentrics inlined. */
typedef struct {
    uint32_t scene;       /* index into a scene table (0 for a single scene) */
    uint32_t pixel;       /* receiver point index */
    uint32_t event;       /* EventDescriptor::asU32; size = event >> 24 */
    uint32_t photon;      /* id of the photon that seeded the tuple */
    int32_t tris[4];
    float params[8];
} palz_path;

/* photonTrace (main_gpu.cu:56-211) over every triangle of the scene,
 * `photons_per_triangle` photons each.  d_out must hold n_tris *
 * photons_per_triangle records.  The traversal stack lives in LDS; a ray that
 * would overflow it sets *d_status (device, may be NULL) to
 * PALZ_ERR_TRAVERSAL_STACK. */
int palz_photon_trace(const palz_scene_view* scene, const palz_photon_params* params,
                      palz_photon_record* d_out, uint32_t* d_status, void* stream);

/* Deterministic replacement of PathCache::store (pk.h:291-349): every photon
 * tuple is inserted into the 2x2x2 cells nearest its landing point; entries
 * are unique per (cell, event, triangles) and keep the smallest photon id
 * (the reference keeps whichever photon arrives first). */
typedef struct palz_tuple_table palz_tuple_table;
int palz_tuple_table_create(palz_tuple_table** out);
int palz_tuple_table_destroy(palz_tuple_table* t);
int palz_tuple_table_build(palz_tuple_table* t, const palz_photon_record* d_records, uint32_t n_records,
                           float cell_step, palz_sorter* sorter, void* stream);
/* Synchronises and reports the table's sizes; any out-pointer may be NULL.
 * Returns PALZ_ERR_CAPACITY if the build overflowed its key layout. */
int palz_tuple_table_info(palz_tuple_table* t, uint32_t* n_valid, uint32_t* n_tuples,
                          uint32_t* n_entries, uint32_t* n_cells);

/* lookupFlatten + sort by path + reorderSpecularPath (main_gpu.cu:373-416,
 * :515-525, pk.h:1486-1541): emit one palz_path per (receiver point, table
 * entry in the point's cell) whose event matches `event` (0 = any event),
 * grouped by tuple.  The count is written to *d_count (device); paths past
 * `capacity` are dropped and the count still reports the full number. */
int palz_lookup_flatten(palz_tuple_table* t, uint32_t event, const palz_first_diffuse* d_points,
                        uint32_t n_points, uint32_t scene, palz_path* d_paths, uint32_t capacity,
                        uint32_t* d_count, palz_sorter* sorter, void* stream);

/* ---- specular path solvers ---------------------------------------------- */

/* Everything a solver needs of one scene (device pointers).  The receiver
 * irradiance of wavelength l at point i accumulates into
 * image[l * n_points + i]: throughput * light_intensity / dAdw * cos, clamped
 * at radiance_clamp (main_gpu.cu:458-483 without the CIE weighting).  The
 * index of refraction is cauchy_a + cauchy_b / lambda^2, lambda in nm
 * (CauchyDispersion, pk.h:935-949). */
typedef struct {
    const palz_triangle* tris;
    const palz_triangle_attrib* attribs;
    const palz_internal_node* nodes;
    const palz_node_index* root;
    const palz_first_diffuse* points;
    const float* lambdas;
    float* image;
    uint32_t n_points;
    uint32_t n_lambdas;
    palz_float3 p_light;
    float light_intensity;   /* m_lightIntencity = 16 (pk.h:1686) */
    float radiance_clamp;    /* m_radianceClamp = 12 (pk.h:1688) */
    float cauchy_a;
    float cauchy_b;
    /* Optional (NULL: the nodes alone): the BVH4 view of the same tree from
     * palz_lbvh_widen, used by the batched solver's occlusion rays. */
    const palz_wide_node* wide;
    /* Optional (NULL: float atomics into `image`): the same sums in unsigned
     * 32.32 fixed point, each contribution truncated to a multiple of 2^-32.
     * The sums are integer additions, so they do not depend on the order in
     * which items finish (the reference's float atomics do); `image` is then
     * not written.  palz_image_fixed_to_float converts. */
    uint64_t* image_fixed;
} palz_scene_desc;

/* image[i] = (float)(fixed[i] * 2^-32), rounded to nearest from the exact
 * double value, for n entries. */
int palz_image_fixed_to_float(const uint64_t* d_fixed, float* d_image, size_t n, void* stream);

/* ---- scene ingestion --------------------------------------------------
 * A Wavefront OBJ file as triangles and attributes, with loadFrame's rules
 * (pk.h:1246-1362; its Alembic input, scene.abc, is not part of the
 * reference): positions and shading normals per face corner, the material
 * named by the `usemtl` in force ("mirror", "dielectric"; any other name or
 * none: `default_material`, a PALZ_MATERIAL_* value; loadFrame's default is
 * Diffuse).  Faces of n > 3 corners become the fan (0, i, i + 1) in file
 * order (loadFrame asserts triangles, pk.h:1320); a corner without a normal
 * takes the triangle's normalised geometric normal.  Indices may be negative
 * (relative).  Host memory only: palz_mesh_copy fills caller buffers of
 * palz_mesh_triangle_count entries (either may be NULL). */
typedef struct palz_mesh palz_mesh;
int palz_mesh_load_obj(const char* path, int default_material, palz_mesh** out);
int palz_mesh_triangle_count(const palz_mesh* m, uint32_t* n);
int palz_mesh_copy(const palz_mesh* m, palz_triangle* tris, palz_triangle_attrib* attribs);
int palz_mesh_destroy(palz_mesh* m);

/* ---- camera side ------------------------------------------------------- */

/* RayGenerator (camera.h:4-41): origin and the image-plane half axes
 * (RayGenerator::lookat fills them from eye, target, up and fovy). */
typedef struct {
    palz_float3 origin;
    palz_float3 right;
    palz_float3 up;
} palz_camera;

/* What solvePrimary (main_gpu.cu:213-371) takes besides the scene.  The lens
 * is LensParams (pk.h:1035-1040) placed in front of the camera (solveLens,
 * pk.h:1042-1073); use_lens = 0 shoots the pinhole rays unrefracted.  The
 * floor texture is an RGBX8 image sampled at (p.x, p.z) (Texture8RGBX::
 * samplePoint, pk.h:1081-1094); without one every diffuse hit has `albedo`. */
typedef struct {
    palz_float3 p_light;
    float light_intensity;   /* 16 (pk.h:1686) */
    float radiance_clamp;    /* 12 (pk.h:1688) */
    float cauchy_a, cauchy_b;
    int32_t iteration;
    uint32_t width, height;
    uint32_t use_lens;
    float lens_distance, lens_thickness, lens_R;
    palz_float3 albedo;
    const uint8_t* texture;  /* device RGBX8, or NULL */
    uint32_t texture_width, texture_height;
} palz_primary_params;

/* solvePrimary (main_gpu.cu:213-371): one jittered camera ray per pixel
 * through the lens to its first diffuse hit (up to 32 specular bounces, a
 * Fresnel choice at dielectrics), a wavelength drawn from the CIE y
 * distribution (CIE_2015_10deg::cmf_y_sample / cmf_y_pdf, pk.h:981-1011),
 * d_first_diffuse[pixel] = {p, ng facing the ray, R, lambda} (only R = 0 is
 * written when no diffuse surface is hit), and d_accum[pixel] += (direct
 * light in linear sRGB / p(lambda), 1).  The exponentials and logarithms are
 * this library's own FP32 routines (the reference uses the hardware's fast
 * ones). */
int palz_solve_primary(const palz_scene_view* scene, const palz_camera* camera, const palz_primary_params* pp,
                       palz_first_diffuse* d_first_diffuse, palz_float4* d_accum, void* stream);

/* Spectral resolve (pk.h:960-1031, main_gpu.cu:9-14, :473-483, :499-513):
 * for each receiver point p, acc = (sum_l (R_p / pi) * xyz2srgblinear(cmf(l)
 * / INTEGRAL_OF_CMF_Y_IN_NM) * d_images[l * n_points + p] * d_weights[l],
 * n_frames) with the CIE 2015 10-degree fit, and pixel = packRGBA(srgb_oetf(
 * acc.xyz / acc.w)).  d_images are the solver's per-wavelength irradiance
 * images (palz_scene_desc.image), d_weights the sample weight of each
 * wavelength (1 / p(lambda) for sampled wavelengths, a quadrature weight for
 * a fixed set).  Either output may be NULL. */
int palz_spectral_resolve(const float* d_images, const float* d_lambdas, const float* d_weights,
                          const palz_first_diffuse* d_points, uint32_t n_points, uint32_t n_lambdas,
                          float n_frames, palz_float4* d_accum, uint32_t* d_pixels, void* stream);

/* Per (path, wavelength) outcome, item index = path * n_lambdas + l. */
enum { PALZ_PATH_NOT_CONVERGED = 0, PALZ_PATH_CONVERGED = 1, PALZ_PATH_CONTRIBUTES = 2, PALZ_PATH_CULLED = 3 };
typedef struct {
    uint32_t status;
    int32_t iterations;   /* Newton iterations run (index of the exit iteration) */
    float contribution;   /* clamped irradiance added to the image */
    float params[8];      /* solved barycentrics */
    float footprint;      /* dA/dw of the contributing chain (pk.h:844-930), else 0 */
} palz_path_result;

/* solveSpecularPath_K{1..4} (main_gpu.cu:418-497): one thread per (path,
 * wavelength).  `event` is the descriptor the path list was flattened for
 * (its size must be K); every path whose own descriptor has K vertices is
 * solved with that descriptor, others are skipped (status NOT_CONVERGED).
 * `d_count` (optional, device) caps n_paths with a device-side count.
 * `d_results` (optional) receives one record per item. */
int palz_solve_specular(int K, uint32_t event, const palz_scene_desc* scene, const palz_path* d_paths,
                        uint32_t n_paths, const uint32_t* d_count, palz_path_result* d_results,
                        void* stream);

/* Per-tuple path-cuts culling: the admissibility chain of
 * traverseAdmissibleNodes at leaf level (pk.h:2101-2207, interval.h:34-713) for
 * every path of chain length 1-3, with the triangles' AABBs as vertex boxes,
 * the AABBs of their shading normals as normal boxes and eta spanning the
 * scene's wavelengths, in interval arithmetic with directed rounding.
 * d_admissible[p] = 0 only if no point of the boxes brings every vertex's
 * Newton residual within the solver's tolerance, so a culled path has no
 * contributing solution; else 1 (also for other chain lengths and for paths
 * naming a scene >= n_scenes).  `d_count` (optional, device) caps n_paths. */
int palz_interval_cull(const palz_scene_desc* d_scenes, uint32_t n_scenes, const palz_path* d_paths,
                       uint32_t n_paths, const uint32_t* d_count, uint8_t* d_admissible, void* stream);
/* The same test in affine arithmetic, after the interval test (no
 * counterpart in the reference, which only sketches affine forms as an
 * idea): each vertex and interpolated shading normal is an affine form in
 * its triangle's two barycentric parameters, and eta one in its own noise
 * symbol, so wi and wo of a vertex keep the vertex they share; coefficients
 * round to nearest with their exact errors (TwoSum, FMA residuals) added
 * outward, square root and reciprocal by linear bounds with directed
 * rounding.  d_admissible[p] = 0 when either test shows the path has no
 * solution; same arguments and guarantees as palz_interval_cull. */
int palz_affine_cull(const palz_scene_desc* d_scenes, uint32_t n_scenes, const palz_path* d_paths, uint32_t n_paths,
                     const uint32_t* d_count, uint8_t* d_admissible, void* stream);

/* Node-level path cuts: traverseAdmissibleNodes<K> (pk.h:1760-2212), called
 * by the reference from host code (main.cpp:930-944, :976-1070), here one
 * walk per receiver point on the device.  For every point of `scene` with a
 * diffuse hit, the K-tuples (K = event >> 24, 1..3) of triangles between
 * scene->p_light and the point that survive the cut traversal over the scene's
 * BVH and normal bounds (palz_lbvh_normal_bounds of the same build), with eta
 * spanning [eta_lo, eta_hi].  The per-tuple test is the residual enclosure of
 * palz_interval_cull, so no tuple with a solution the solvers accept is cut.
 * Output: palz_path records grouped by point in point order, each point's
 * tuples in the reference's traversal order (LIFO, child 1 first), with
 * photon = 0xFFFFFFFF, `scene_index` as scene, and 1/3 as every start
 * barycentric (pk.h:470-473).  *d_count (device) receives the total; records
 * past `capacity` are dropped (capacity 0: count only).  A walk deeper than
 * its stack (192 entries) stops early and sets *d_status (device, may be
 * NULL) to PALZ_ERR_TRAVERSAL_STACK; a total of 2^32 tuples or more sets
 * PALZ_ERR_CAPACITY, writes 0xFFFFFFFF as the count and no records.  At most
 * 2^24 points per call.  The call synchronises `stream` once per level of its
 * breadth-first expansion.  `scene` is a host struct holding device pointers;
 * its points, light, triangles and BVH are read. */
typedef struct palz_path_cuts palz_path_cuts;
int palz_path_cuts_create(palz_path_cuts** out);
int palz_path_cuts_destroy(palz_path_cuts* pc);
int palz_path_cuts_enumerate(palz_path_cuts* pc, const palz_scene_desc* scene, const palz_normal_bound* d_normal_bounds,
                             uint32_t event, float eta_lo, float eta_hi, uint32_t scene_index, palz_path* d_paths,
                             uint32_t capacity, uint32_t* d_count, uint32_t* d_status, void* stream);

/* ---- renderer: a PKRenderer-shaped host driver (pk.h:1363-1640) ----------
 * Owns the scene, BVH, photon records, tuple table, path list, batch and the
 * per-pixel buffers of a W x H image, and calls the entry points above in the
 * reference's order.  palz_renderer_load_scene / _load_obj = loadFrame (host
 * arrays in; BVH built); palz_renderer_step = step: solvePrimary, photonTrace
 * over the eta range of the wavelengths, the tuple table, one lookup of every
 * chain of length 1-3 and the batched solve (the reference runs one lookup
 * and launch per descriptor), then the T,T,T,T chains with the plain solver
 * (the reference's K = 4 launch, pk.h:1563), and the frame's per-wavelength
 * images (fixed-point sums: independent of the solvers' item order) resolved (palz_spectral_resolve: CIE fit, `weights`, the frame's
 * receiver reflectance) into the float4 accumulators; palz_renderer_clear =
 * clear; palz_renderer_resolve = resolve (pack) into host RGBA8.  The
 * reference draws one wavelength per pixel and frame; here every frame
 * solves the fixed list `lambdas` (nm) with quadrature `weights`.  Paths past
 * path_capacity in a frame are dropped (counts: palz_renderer_buffers). */
typedef struct {
    palz_float3 p_light;            /* m_p_light */
    float light_intensity;          /* 16 (pk.h:1686) */
    float radiance_clamp;           /* 12 (pk.h:1688) */
    float min_throughput;           /* 0.05 (pk.h:1687) */
    float cauchy_a, cauchy_b;       /* CauchyDispersion, lambda in nm */
    uint32_t photons_per_triangle;  /* 256 (pk.h:1456-1457) */
    uint32_t use_lens;
    float lens_distance, lens_thickness, lens_R; /* LensParams (pk.h:1690-1694) */
    palz_float3 albedo;             /* receiver reflectance */
    uint32_t path_capacity;         /* paths per frame */
    int32_t refine;                 /* PALZ_REFINE_* of the batched solve */
    uint32_t solve_tttt;            /* 1: solve T,T,T,T chains too */
} palz_renderer_params;

typedef struct palz_renderer palz_renderer;
int palz_renderer_create(uint32_t width, uint32_t height, const float* lambdas, const float* weights,
                         uint32_t n_lambdas, palz_renderer** out);
int palz_renderer_destroy(palz_renderer* r);
int palz_renderer_set_params(palz_renderer* r, const palz_renderer_params* p);
int palz_renderer_get_params(const palz_renderer* r, palz_renderer_params* p);
int palz_renderer_set_camera(palz_renderer* r, const palz_camera* c);
int palz_renderer_load_scene(palz_renderer* r, const palz_triangle* tris, const palz_triangle_attrib* attribs,
                             uint32_t n, void* stream);
int palz_renderer_load_obj(palz_renderer* r, const char* path, int default_material, void* stream);
int palz_renderer_step(palz_renderer* r, void* stream);
int palz_renderer_clear(palz_renderer* r, void* stream);
int palz_renderer_resolve(palz_renderer* r, uint32_t* rgba8, void* stream);
/* Device views: the last frame's first-diffuse records, the accumulators and
 * three counts (photon trace status, paths of length 1-3, T,T,T,T paths). */
int palz_renderer_buffers(const palz_renderer* r, const palz_first_diffuse** d_points, const palz_float4** d_accum,
                          const uint32_t** d_counts);

/* ---- adjoint of receiver irradiance (no reference counterpart) -----------
 * Heights of a heightfield as parameters: a triangle's vertex c has height
 * parameter tri_vertex[3 t + c] (-1: the triangle is not part of the
 * heightfield); vertex y = height, and the vertex's shading normal is
 * normalize(-dh/dx, 1, -dh/dz) with the derivatives taken as central (one-sided
 * at the border) differences: normal_nb[4 h + j] names the j-th height the
 * normal of parameter h depends on (-1: none) and normal_coef[8 h + 2 j + 0/1]
 * its coefficients in dh/dx and dh/dz.  All three arrays are device memory. */
typedef struct {
    const int32_t* tri_vertex;
    const int32_t* normal_nb;
    const float* normal_coef;
    uint32_t n_params;
    uint32_t pad_;
} palz_height_param;

/* d_grad[h] += sum over contributing items of d_weight[l * n_points + pixel]
 * * d(irradiance contribution)/d(height h), through the converged Newton
 * solution by the implicit-function theorem (FP64; the tuple set, visibility
 * and validity of the forward solve are held).  d_weight is dL/dI of the
 * receiver image (n_lambdas x n_points).  With d_results (the forward solve's
 * per-item records, d_paths in the same order) the contributing items and
 * their FP32 solutions are read; with d_results NULL the forward solve is
 * recomputed per item instead of stored. */
int palz_adjoint_irradiance(const palz_scene_desc* scene, const palz_height_param* hp, const palz_path* d_paths,
                            uint32_t n_paths, const palz_path_result* d_results, const float* d_weight,
                            double* d_grad, void* stream);

/* ---- batched persistent solver (no reference counterpart) ---------------
 * Replaces the reference's per-event-descriptor launch loop (pk.h:1480-1586,
 * one lookupFlatten + blocking count read + solveSpecularPath_K launch per
 * descriptor, per scene) with one batch over many scenes and chain lengths.
 * Work items are (path, wavelength) with item = path * n_lambdas + l.  For
 * each K = 1..3 a persistent kernel refills idle lanes at Newton-iteration
 * granularity from a chunked work queue; converged items then go through the
 * throughput / dA/dw stage and, with `refine`, one FP64 Newton step on the
 * converged solution (PALZ_REFINE_*).  Chains with K = 4 are not
 * part of a batch (status NOT_CONVERGED, counted in stats[3]).  Every count
 * stays on the device, so a solve can be captured in a hipGraph. */
typedef struct palz_batched palz_batched;

/* Launch shape of one chain length: lanes per block, resident blocks per
 * CU (grid = blocks_per_cu * CU count), items handed out per queue ticket. */
typedef struct {
    int32_t block;
    int32_t blocks_per_cu;
    int32_t chunk;
    int32_t finish_block;
} palz_launch_config;

int palz_batched_create(palz_batched** out);
int palz_batched_destroy(palz_batched* b);
/* Sorts n_paths paths into the batch's own buffer by chain length, then by
 * scene, then by first triangle (its index shifted down so that the key fits
 * 32 bits), keeping the input order of equal keys: each chain length's items
 * come grouped by scene and by region of the scene, so that waves hold one
 * scene's neighbouring chains and each XCD's share of a list covers one part
 * of a scene.  `d_count` (optional, device) replaces n_paths, which is then
 * the capacity.  After the call, palz_batched_order gives, for every batch
 * slot, the index of the path in d_paths. */
int palz_batched_prepare(palz_batched* b, const palz_path* d_paths, uint32_t n_paths, const uint32_t* d_count,
                         palz_sorter* sorter, void* stream);
int palz_batched_order(palz_batched* b, const uint32_t** d_order, const palz_path** d_sorted_paths,
                       const uint32_t** d_ranges);
/* K = 1..3: the kernels of chain length K.  K = 0: the mixed-length kernels
 * of fused mode: block and blocks_per_cu of the Newton kernel, finish_block
 * of the occlusion kernel, and in `chunk` the Newton kernel's register budget
 * in waves per SIMD, 2 or 3 (each length keeps its own queue chunk). */
int palz_batched_set_config(palz_batched* b, int K, const palz_launch_config* cfg);
int palz_batched_get_config(palz_batched* b, int K, palz_launch_config* cfg);
/* Fused mode (no counterpart in the reference), a bit mask: 1 = one
 * persistent Newton kernel over every chain length (each wave drains the
 * K = 3 queue, then K = 2's, then K = 1's) in place of one per length; 2 = one
 * occlusion kernel over every chain length likewise; 3 = both; 0 = one of
 * each per chain length.  The finish kernels stay per length.  -1 (default)
 * = per length until palz_batched_autotune has timed whole solves in all four
 * structures and kept the fastest.  Results are the same in every structure.
 * Drops the captured graph. */
int palz_batched_set_fused(palz_batched* b, int mode);
/* The fused mode set and the mask solves now run with (0..3). */
int palz_batched_get_fused(palz_batched* b, int* mode, int* active);
/* Solve every prepared item.  d_scenes: device array of n_scenes scene
 * descriptors (a path's `scene` indexes it); every scene's lambdas array
 * holds at least n_lambdas entries.  d_results (optional) is indexed by batch
 * item (batch slot * n_lambdas + l); d_refined (optional) receives 8 doubles
 * per item: the FP64-refined barycentrics of converged items. */
enum {
    PALZ_REFINE_NONE = 0,     /* FP32 Newton only: results equal palz_solve_specular's */
    PALZ_REFINE_VERTICES = 1, /* + one FP64 Newton step on converged chains; reported
                                 params are refined, irradiance uses the FP32 chain */
    PALZ_REFINE_ALL = 2       /* as 1, and the refined chain also feeds dA/dw */
};
int palz_batched_solve(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t n_lambdas,
                       int refine, palz_path_result* d_results, double* d_refined, void* stream);
/* Capture palz_batched_solve with these arguments in a hipGraph (first call
 * or changed arguments; captured on the batch's own stream) and launch the
 * graph on `stream`.  palz_batched_prepare and palz_batched_set_config drop
 * the captured graph. */
int palz_batched_solve_graph(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes,
                             uint32_t n_lambdas, int refine, palz_path_result* d_results, double* d_refined,
                             void* stream);
/* How a captured solve is laid out (no counterpart in the reference): 1 =
 * three graph branches, one per chain length (forked after the reset, joined
 * for the compaction, forked again for the finishing kernels), so that one
 * length's short kernels overlap the others'; 0 = one stream; -1 (default) =
 * branches when no chain length holds three quarters of the paths (read once,
 * at capture).  Drops the captured graph. */
int palz_batched_set_branches(palz_batched* b, int mode);
/* Per chain length K = 1..3 (row K-1): [0] items, [1] converged, [2]
 * contributing, [3] residual evaluations (Newton iterations plus the final
 * converged check); row 3: [0] items culled, [2] items whose path names a
 * scene >= n_scenes (not solved, status NOT_CONVERGED), [3] items skipped
 * (K = 4).
 * Synchronises the stream. */
int palz_batched_stats(palz_batched* b, uint64_t* stats16, void* stream);
/* Per chain length K = 1..3: the number of paths of the last solve with at
 * least one converged (path, wavelength) item, i.e. the paths whose records,
 * triangles and receiver point the finishing kernels read.  No counterpart in
 * the reference (its solver keeps no statistics).  Synchronises the stream. */
int palz_batched_path_stats(palz_batched* b, uint64_t* converged_paths3, void* stream);
/* Kernel timing of eager solves (graph launches are not timed): when
 * enabled, each palz_batched_solve records HIP events on its stream around
 * its kernels; palz_batched_timings waits for the last solve's events and
 * returns 12 durations in ms: Newton K=1, Newton K=2, Newton K=3, the
 * compaction of the converged items, then for K = 1, 2, 3 the occlusion rays
 * (contributableThroughput) and the finish (dA/dw, irradiance, results and,
 * with refinement on, the FP64 refinement step), then a refinement slot that
 * stays ~0 (the step runs inside the finish kernels) and the interval culling
 * (when enabled; else ~0). */
int palz_batched_set_timing(palz_batched* b, int enable);
int palz_batched_timings(palz_batched* b, float* ms12);
/* Per-tuple culling in front of the Newton kernels: items of a culled path
 * get status PALZ_PATH_CULLED without a solve.  mode 1: the interval test
 * (palz_interval_cull); 2: the interval test, then the affine test
 * (palz_affine_cull) on the paths it keeps; 0 (default): off.  A culled path
 * has no contributing solution, so the image and the contributing items are
 * those of an unculled solve. */
int palz_batched_set_culling(palz_batched* b, int mode);
/* Time every candidate launch shape per chain length on the prepared batch
 * and keep the fastest (the autotuner).  The image receives the candidates'
 * contributions; clear it afterwards.  Writes the chosen shapes and their
 * times to `report` (3 rows of {block, blocks_per_cu, chunk, finish_block,
 * newton_ms * 1000, finish_ms * 1000}) if not NULL. */
int palz_batched_autotune(palz_batched* b, const palz_scene_desc* d_scenes, uint32_t n_scenes, uint32_t n_lambdas,
                          int refine, int32_t* report, void* stream);
/* The autotuner's tile-shape cache (no counterpart in the reference): with a
 * file set, palz_batched_autotune first looks the workload up there (key: CU
 * count, scene count and per-length item counts rounded up to powers of two,
 * n_lambdas, refine, culling and fused modes) and applies the stored shapes
 * (and, in fused mode -1, the stored choice) without timing anything;
 * otherwise it tunes and appends one line.  NULL turns the cache off (the
 * default). */
int palz_batched_set_tune_cache(palz_batched* b, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* PALZ_H */
