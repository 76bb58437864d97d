/*
 * cvt.h -- C ABI of the MI355X CVT engine (libcvt_hip.so).
 *
 * Plain pointers and sizes only.  Every pointer argument may be a host
 * pointer or a device pointer of the engine's device (copies use
 * hipMemcpyDefault); calls that read results back to a host pointer
 * synchronise the engine's stream.  Every function returns CVT_OK (0) or a
 * negative status; cvt_last_error() gives the message of the calling
 * thread's last failure.  A context is not re-entrant: one thread at a time.
 * The one-shot calls without a context (cvt_rasterize, cvt_extract_points,
 * cvt_locate_points, cvt_vertex_weights) run on the legacy default stream,
 * as the reference's calls do: they wait for work queued there and on
 * blocking streams, and return when their results are written.
 *
 * Conventions: a grid of width x height pixels, row-major (pixel (x, y) at
 * y*width + x, the reference's TOID, source/gcvt.h:26).  Sites are fp32
 * (x, y) pairs in pixel units; pixel (x, y) has its centre at
 * (x + 0.5, y + 0.5).  Labels are int32 site indices, CVT_LABEL_NONE where no
 * site reaches.  The Voronoi map is JFA with steps s0, s0/2, ..., 1 plus one
 * extra step-1 pass (s0 = largest power of two below max(width, height)),
 * distance d = fmaf(dy, dy, dx*dx) (dx*dx rounded first), ties to the
 * lower index.
 *
 * Which reference interface each entry point replaces:
 *   cvt_create / cvt_destroy      gcvtInitialization / pbaCVDDeinitialization
 *                                 (source/gcvt.cu:840-903)
 *   cvt_set_density               pba2DInitializeInput density upload and the
 *                                 once-per-run prefix sums
 *                                 (source/gcvt.cu:906-914, :995-1006)
 *   cvt_set_sites / cvt_get_sites the site texture upload / download
 *                                 (source/gcvt.cu:1101-1103, :1152)
 *   cvt_voronoi / cvt_get_labels  pba2DCompute (source/gcvt.cu:967-978)
 *   cvt_moments                   pbaCVDComputeCentroid
 *                                 (source/gcvt.cu:1008-1023)
 *   cvt_lloyd                     the Lloyd loop body: pba2DCompute +
 *                                 pbaCVDCalcEnergy + pbaCVDComputeCentroid +
 *                                 pbaCVDUpdateSites
 *                                 (source/gcvt.cu:1113-1121)
 *   cvt_lloyd_until               the same loop with its stop rule
 *                                 (source/gcvt.cu:1123-1140), tested on the
 *                                 device
 *   cvt_energy                    pbaCVDCalcEnergy (source/gcvt.cu:1059-1083),
 *                                 in float64 and without its reduction bias
 *   cvt_compat_gcvt               gCVT (source/gcvt.cu:1087-1156)
 *   cvt_rasterize                 discretization_d
 *                                 (source/discretization.cu:87-120)
 *   cvt_extract_points            the site scan of delaunayInput
 *                                 (source/delaunay.h:48-58)
 *   cvt_locate_points             locate() and the vertex interpolation of
 *                                 recover() (source/recover.h:63-81,
 *                                 :109-125, :134-141)
 *   cvt_vertex_weights            weighting_vertices_with_area
 *                                 (source/weighting.h:37-97)
 */
#ifndef CVT_H
#define CVT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVT_OK 0
#define CVT_ERR_ARG (-1)      /* invalid argument */
#define CVT_ERR_HIP (-2)      /* HIP runtime failure */
#define CVT_ERR_NOMEM (-3)    /* device allocation failed */
#define CVT_ERR_STATE (-4)    /* call out of order (e.g. no density yet) */
#define CVT_ERR_CAPACITY (-5) /* a fixed-size table overflowed */

#define CVT_LABEL_NONE 0x7fffffff

typedef struct cvt_ctx cvt_ctx;

/* Message of the calling thread's last failure ("" if none). */
const char *cvt_last_error(void);
/* ABI version (major * 100 + minor). */
int cvt_version(void);

/* A context for one grid of at most 8192 x 8192 pixels (the reference's
 * limit, source/gcvt.cu:842-847): allocates every device buffer it will use
 * (the reference allocates per call, source/gcvt.cu:859-874).  max_sites
 * bounds the site count of later calls. */
int cvt_create(cvt_ctx **ctx, int width, int height, int max_sites);
int cvt_destroy(cvt_ctx *ctx);
/* Launch subsequent work on this hipStream_t (NULL = the context's own
 * non-blocking stream). */
int cvt_set_stream(cvt_ctx *ctx, void *hip_stream);
/* width*height fp32 densities, finite and >= 0. */
int cvt_set_density(cvt_ctx *ctx, const float *density);
/* num_sites (x, y) fp32 pairs in pixel units, finite and inside
 * [-width, 2 width] x [-height, 2 height]; fixed (optional, num_sites bytes)
 * marks sites the update never moves (the reference's constraint mask,
 * source/gcvt.h:106-122). */
int cvt_set_sites(cvt_ctx *ctx, const float *xy, const uint8_t *fixed, int num_sites);
int cvt_get_sites(cvt_ctx *ctx, float *xy);
/* Current site count (>= 0) or a negative status. */
int cvt_num_sites(cvt_ctx *ctx);
/* Label map of the current sites. */
int cvt_voronoi(cvt_ctx *ctx);
/* width*height int32 labels of the last Voronoi pass. */
int cvt_get_labels(cvt_ctx *ctx, int32_t *labels);
/* Per-site float64 moments over the cells of the last Voronoi pass, 4 per
 * site: W = sum rho, DX = sum rho (px - sx), DY = sum rho (py - sy),
 * E = sum rho |p - s|^2 (pixel units); the centroid is s + (DX, DY) / W. */
int cvt_moments(cvt_ctx *ctx, double *moments);
/* iterations Lloyd steps: Voronoi of the current sites, moments, energy,
 * then s <- s + omega (c - s) for every non-fixed site with W > 0.  The
 * per-iteration energy and gradient norm are recorded (cvt_get_history). */
int cvt_lloyd(cvt_ctx *ctx, int iterations, float omega);
/* Lloyd steps as cvt_lloyd, at most max_iter, that stop themselves on the
 * device: after the first iteration whose energy E is <= target, or whose
 * energy fell by less than rel_tol * E over the last `window` iterations of
 * this call (window 0: no such test; target -inf: no target).  The update's
 * last block tests the rule on the device history and raises a stop flag
 * that makes every kernel of the later iterations return at entry.  The
 * host enqueues chunks of 8 iterations, each followed by a copy of the flag,
 * and waits only for the chunk before the one still queued, so the device
 * never idles and at most two chunks run past the stop (as no-ops).
 * *iterations (optional) receives the number of iterations run; the history
 * holds their energies as with cvt_lloyd.  The reference tests its stop rule
 * on the host every 10 iterations after a blocking copy of the energy
 * (source/gcvt.cu:1123-1140). */
int cvt_lloyd_until(cvt_ctx *ctx, float omega, double rel_tol, int window, double target,
                    int max_iter, int *iterations);
/* Number of recorded iterations (>= 0) or a negative status. */
int cvt_history_length(cvt_ctx *ctx);
/* The first n recorded (energy, gradient norm) pairs, 2n doubles.  Energy
 * E = sum_p rho_p |s(p) - p|^2 / (width * height) in pixel units (the
 * reference's normalisation, source/gcvt.cu:788-802); |grad E| =
 * 2 / (width * height) * sqrt(sum_i DX_i^2 + DY_i^2). */
int cvt_get_history(cvt_ctx *ctx, double *energy_gradnorm, int n);
int cvt_clear_history(cvt_ctx *ctx);
/* Float64 energy of the current sites over their fresh Voronoi map, and the
 * gradient norm, reduced on the GPU in a fixed order. */
int cvt_energy(cvt_ctx *ctx, double *energy, double *gradnorm);
int cvt_synchronize(cvt_ctx *ctx);
/* Runs `iterations` (1..1000) Lloyd steps eagerly with HIP events around
 * each phase and returns mean durations in ms: ms[0] one dense JFA pass (the
 * back-to-back sequence of passes 1 .. P-1 over P - 1), ms[1] the moment
 * kernel, ms[2] the first JFA pass (from the seed map), ms[3] update +
 * finalize, ms[4] the whole JFA, ms[5] the whole step,
 * ms[6] the JFA pass count P, ms[7] the moment tile.
 * (No reference counterpart: the reference times stages with clock(),
 * source/main.cpp:181-242.) */
int cvt_profile(cvt_ctx *ctx, int iterations, float omega, double *ms);

/* Reference-compatible integer-site CVT (the semantics of gCVT,
 * source/gcvt.cu:1087-1156): voronoi is the size x size map of interleaved
 * (x, y) int16 labels (MARKER = -32768), read for the initial sites (pixels
 * labelled with their own coordinates) and overwritten with the final label
 * map; density is size x size fp32; mask (optional, one byte per pixel)
 * marks fixed sites.  depth > 1 adds coarse levels down to 256 on a 2x2-mean
 * density pyramid, as the reference does (source/gcvt.cu:1087-1147): the
 * initial sites are read from the first nc^2 (x, y) pairs of voronoi as an
 * nc x nc map (nc = the coarsest side), the mask is read at each level with
 * that level's row stride, the iteration count, omega and the energies of the
 * decade checks run on across levels, and each level runs at least one
 * iteration (so the count can exceed max_iter by depth - 1).
 * iterations / num_sites (optional) receive the iteration count and the final
 * site count.  Differences from the reference: JFA(+1) instead of exact PBA+
 * for the Voronoi map; float64 moments and energy; a pixel is a site only
 * when it holds its own coordinates (PBA+ also takes a pixel holding other
 * coordinates as a site, and in the coarse map read from a full-size one
 * such values can lie outside the coarse grid, where the reference's
 * kernelVoronoi1D writes outside its shared array); any size in [2, 8192]
 * (the reference needs a power of two >= 256). */
int cvt_compat_gcvt(int16_t *voronoi, const float *density, const uint8_t *mask, int size,
                    int depth, int max_iter, int *iterations, int *num_sites);

/* The reference's initial sites for gCVT (putConstrains + randomPoints,
 * source/gcvt.h:76-122, run by centroidalVoronoi at :124-139), as a call a
 * Linux host can make: the reference's gcvt.h does not compile on glibc
 * (its `double random()` collides with POSIX `long random()`).  voronoi
 * (size x size interleaved (x, y) int16) is overwritten: mask pixels (mask
 * may be null) hold their own coordinates, every other pixel MARKER; then
 * num sites are drawn by rejection (x = floor(u size), y = floor(u size),
 * z = u mx, accepted when the pixel is free and density > z, with
 * mx = min(max density, 100 * mean of the non-zero density)).  u is the KISS
 * generator of source/gcvt.h:58-74 with the reference's LLP64 semantics
 * (32-bit words, u = draw / 2^32).  rng_state: the generator's four words
 * (z, w, jsr, jcong), read and advanced; NULL starts from four zero words,
 * as the reference does (it never calls randinit).  draws (optional)
 * receives the number of generator draws.  Runs on the host (the loop is
 * sequential); buffers may be host or device memory.  CVT_ERR_ARG when
 * fewer free pixels of positive density exist than num (the reference's loop
 * would not end) or the density holds a negative, NaN or infinite value. */
int cvt_ref_init_sites(int16_t *voronoi, const float *density, const uint8_t *mask, int size,
                       int num, uint32_t *rng_state, long long *draws);

/* Density rasterizer (the semantics of discretization_d,
 * source/discretization.cu:57-120): density[y*n + x] for the sample point
 * (x*scale, y*scale) is the float64 barycentric interpolation of weight over
 * the first triangle (index order) containing it, 0 outside every triangle.
 * points holds num_point (x, y) pairs already shifted to the origin,
 * triangle 3 vertex indices per triangle. */
int cvt_rasterize(const double *points, const double *weight, int num_point, const int *triangle,
                  int num_tri, float *density, double scale, int n);
/* cvt_rasterize for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc); it
 * has no counterpart in the reference.  Same arguments and the same result
 * bit for bit, but every buffer must be device-accessible (device, pinned or
 * managed memory; CVT_ERR_ARG otherwise), and the call allocates nothing:
 * its working memory is a static device array of about 14 MB that each
 * device holds until the process ends.  Limits: n <= 4096, num_tri <= 262144
 * and at most 2097152 (triangle, 16x16-pixel tile) pairs (CVT_ERR_ARG /
 * CVT_ERR_CAPACITY beyond).  Calls that overlap on one device share that
 * array and race; like cvt_rasterize it runs on the legacy default stream
 * and returns when the density is written. */
int cvt_rasterize_noalloc(const double *points, const double *weight, int num_point,
                          const int *triangle, int num_tri, float *density, double scale, int n);

/* CDT input points (the site scan of delaunayInput, source/delaunay.h:48-58):
 * pixels (x, y) whose label is their own coordinates and whose mask byte is
 * 0 (mask may be null), in the reference's order (x outer, y inner), as
 * (x*scale + l, y*scale + b) float64 pairs.  *count receives the number
 * found; at most capacity pairs are written (CVT_ERR_CAPACITY if more);
 * out = NULL only counts. */
int cvt_extract_points(const int16_t *voronoi, const uint8_t *mask, int n, double scale, double l,
                       double b, double *out, long capacity, long *count);

/* Point location in a 2D triangle mesh (replaces the brute-force face scan
 * locate() of recover(), source/recover.h:63-81, run for every site at
 * :109-125 and for every output triangle's centroid at :134-141): for each
 * query point q[k] the first face, in face order, whose float64 barycentric
 * weights (the formula of source/recover.h:30-53) are all >= 0, or -1 for
 * none (a face with a zero denominator holds no point; a non-finite point is
 * in no face).  uv: [nv][2] vertex coordinates; tri: [nf][3] vertex indices
 * (CVT_ERR_ARG if one is outside [0, nv)); q: [nq][2]; face: [nq].
 * Optional: w [nq][3] receives the weights of the face's vertices 0, 1, 2
 * (1 - w1 - w2, w1, w2 as recover() stores them; -1 where face is -1); with
 * p3 [nv][3] (the 3D vertices), out3 [nq][3] receives sum_j w_j p3[v_j]
 * summed in j order from 0.0 (source/recover.h:117-123; 0 where face is -1).
 * Host or device pointers; the call returns when the results are written. */
int cvt_locate_points(const double *uv, const double *p3, int nv, const int *tri, int nf,
                      const double *q, int nq, int *face, double *w, double *out3);

/* Per-vertex density weights of a parameterized mesh, the rasterizer's vertex
 * weights (replaces weighting_vertices_with_area, source/weighting.h:37-97):
 * weight[v] = (sum of the 3D areas of v's faces) / (sum of their 2D areas),
 * each sum taken in face order; a weight above 1e10 is replaced by the
 * largest other weight (0 if there is none); a vertex in no face gets NaN.
 * uv: [nv][2] 2D vertices; p3: [nv][3] the 3D position of each 2D vertex;
 * tri: [nf][3] vertex indices in vertices_around_face order (CVT_ERR_ARG if
 * one is outside [0, nv)); weight: [nv].  The 2D area is get_area
 * (source/weighting.h:29-35); the 3D area is CGAL's PMP::face_area on the
 * points (tri[2], tri[0], tri[1]).  Host or device pointers; the call
 * returns when weight is written. */
int cvt_vertex_weights(const double *uv, const double *p3, int nv, const int *tri, int nf,
                       double *weight);

/* ---------------------------------------------------------------------------
 * Batched CVT: many independent patches of different sizes advanced
 * together, each launch covering every patch (no reference counterpart: the
 * reference runs one mesh per gCVT call).  A patch's results equal a
 * single-patch context's on the same inputs (labels bit for bit; moments bit
 * for bit at the same moment tile).  Patch widths must be multiples of 4 in
 * [8, 8192].
 * ------------------------------------------------------------------------- */
typedef struct cvt_batch cvt_batch;

int cvt_batch_create(cvt_batch **batch, int num_patches, const int *widths, const int *heights,
                     const int *max_sites);
int cvt_batch_destroy(cvt_batch *batch);
int cvt_batch_set_stream(cvt_batch *batch, void *hip_stream);
int cvt_batch_set_density(cvt_batch *batch, int patch, const float *density);
int cvt_batch_set_sites(cvt_batch *batch, int patch, const float *xy, const uint8_t *fixed,
                        int num_sites);
int cvt_batch_get_sites(cvt_batch *batch, int patch, float *xy);
/* iterations Lloyd steps of every patch (one captured graph per step). */
int cvt_batch_lloyd(cvt_batch *batch, int iterations, float omega);
/* Label maps of every patch for their current sites. */
int cvt_batch_voronoi(cvt_batch *batch);
int cvt_batch_get_labels(cvt_batch *batch, int patch, int32_t *labels);
/* Recorded (energy, gradient norm) pairs of a patch: copies min(n, length)
 * pairs and returns that count (n = 0 returns the length). */
int cvt_batch_history(cvt_batch *batch, int patch, double *energy_gradnorm, int n);
/* Moment tile of a patch (64, 128 or 256): results do not depend on it. */
int cvt_batch_set_tile(cvt_batch *batch, int patch, int tile);
int cvt_batch_get_tile(cvt_batch *batch, int patch);
int cvt_batch_synchronize(cvt_batch *batch);

/* ---------------------------------------------------------------------------
 * Learned warm start (no reference counterpart): a 4-level convolutional
 * U-Net, 32 / 64 / 96 / 128 channels, BF16 weights and activations with FP32
 * accumulation on MFMA, FP32 master weights, Adam.  Input R x R (R = 256 for
 * the warm start): density box mean / mean, initial-site count per cell *
 * R^2 / S averaged over a (2h+1)^2 window with h = round(R / sqrt(S)) - 1,
 * and the domain indicator; output a 2-channel displacement field
 * (units of the domain side) sampled bilinearly at the sites.  All buffers
 * may be host or device pointers.
 * ------------------------------------------------------------------------- */
/* ---------------------------------------------------------------------
 * Delaunay triangulation of 2D points (the unconstrained part of gDel2D's
 * GpuDel::compute, source/gDel2D/gDel2D/GPU/GpuDelaunay.cu:46-83, with
 * GDel2DInput / GDel2DOutput at source/gDel2D/gDel2D/GpuDelaunay.h:28-64).
 * A context owns the device memory for up to max_points points.
 * cvt_dt_compute takes num_points float64 (x, y) pairs and writes the
 * finite Delaunay triangles (counterclockwise vertex index triples, like
 * GDel2DOutput::triVec) and, if opp is not null, their neighbours in
 * gDel2D's TriOpp encoding ((triangle << 4) | edge, edge i opposite vertex
 * i; -1 on the convex hull; GDel2DOutput::triOppVec).  *num_tri receives the
 * count (CVT_ERR_CAPACITY when it exceeds capacity triangles).  Predicates
 * are exact, and exact degeneracies (collinear, cocircular points) are
 * decided by a symbolic perturbation of the points by index, as gDel2D's
 * SoS predicates do: the result is a Delaunay triangulation holding every
 * distinct point, with zero-area triangles along collinear hull points; in
 * general position (no three points collinear, no four cocircular) it is
 * the unique one.  Exact duplicate points are dropped (cvt_dt_stats counts
 * them).  CVT_ERR_ARG for non-finite points, for inputs whose points are
 * all collinear, and for coordinates whose magnitudes span more than 2^71
 * units of their finest bit.  Arrays may be host or device memory; the
 * call returns when the output is written.
 * --------------------------------------------------------------------- */
typedef struct cvt_dt cvt_dt;
int cvt_dt_create(cvt_dt **out, int max_points);
int cvt_dt_compute(cvt_dt *dt, const double *points, int num_points, int *tri, int *opp,
                   int capacity, int *num_tri);
/* The same with constraint segments (GDel2DInput::constraintVec; gDel2D's
 * doInsertConstraints, GpuDelaunay.cu:638-747): num_segments pairs of point
 * indices that become edges of the result, a constrained Delaunay
 * triangulation (every other edge has an empty circle among the points
 * visible across it).  Their links in opp carry bit 2, as gDel2D's TriOpp
 * marks constraints.  Segments may share end points; segments that cross
 * each other or pass through a point give CVT_ERR_ARG. */
int cvt_dt_compute_constrained(cvt_dt *dt, const double *points, int num_points,
                               const int *segments, int num_segments, int *tri, int *opp,
                               int capacity, int *num_tri);
/* stats[4]: insertion rounds, flip passes, flips and dropped duplicates of
 * the last cvt_dt_compute (gDel2D's Statistics, CommonTypes.h:532) */
int cvt_dt_stats(cvt_dt *dt, int *stats);
int cvt_dt_destroy(cvt_dt *dt);
/* orient[i] = sign of orient2d(a, b, c) and incircle[i] = sign of
 * incircle(a, b, c, d) for count 8-double records (ax ay bx by cx cy dx dy),
 * evaluated on the device by the triangulation's exact predicates (gDel2D's
 * orient2dFastExact / incircleExact, source/gDel2D/gDel2D/GPU/KerShewchuk.h:636,726) */
int cvt_dt_predicates(const double *quads, int count, int *orient, int *incircle);

typedef struct cvt_unet cvt_unet;

int cvt_unet_create(cvt_unet **net, int R, int max_batch, unsigned long long seed);
int cvt_unet_destroy(cvt_unet *net);
int cvt_unet_set_stream(cvt_unet *net, void *hip_stream);
long cvt_unet_num_params(cvt_unet *net);
int cvt_unet_get_params(cvt_unet *net, float *params);
int cvt_unet_set_params(cvt_unet *net, const float *params);
int cvt_unet_get_grads(cvt_unet *net, float *grads);
/* Input of batch slot `slot` (dst = NULL) or into dst (R*R*32 bf16) from an
 * n x n density (n a multiple of R) and S sites in pixel units. */
int cvt_unet_prepare(cvt_unet *net, int slot, const float *density, int n, const float *xy, int S,
                     uint16_t *dst);
/* nb inputs (nb * R * R * 32 bf16, NHWC) into the batch slots. */
int cvt_unet_set_input(cvt_unet *net, const uint16_t *x, int nb);
int cvt_unet_forward(cvt_unet *net, int nb);
int cvt_unet_get_field(cvt_unet *net, float *field, int nb);
/* Loss (mean |field(uv) - target|^2 over nb * S sites, uv = position / side)
 * and the backward pass into the gradients; *loss (optional) synchronises. */
int cvt_unet_loss_backward(cvt_unet *net, int nb, const float *uv, const float *target, int S,
                           float *loss);
int cvt_unet_adam(cvt_unet *net, float lr, float beta1, float beta2, float eps);
/* Displacements of S sites of batch slot `slot` from the last forward pass. */
int cvt_unet_sample(cvt_unet *net, int slot, const float *uv, int S, float *disp);
/* Warm start of an engine context in place, on the context's stream and
 * without host round trips: the input is built from the context's density
 * and current sites (the grid must be square, a multiple of R), the network
 * runs, and each free site moves by the field sampled at its position.  A
 * site keeps its position when it is fixed or when the target pixel lies off
 * the grid or has zero density.  Uses batch slot 0.  Plain Lloyd continues
 * from the moved sites; the caller keeps the old sites to fall back to.
 * A network object serves one call at a time (its input buffers are
 * shared): concurrent warm starts need one network each. */
int cvt_unet_warm_start(cvt_unet *net, cvt_ctx *ctx);
/* Weights as a raw bf16 blob ("CVTUNET1", R, count, values). */
int cvt_unet_save(cvt_unet *net, const char *path);
int cvt_unet_load(cvt_unet *net, const char *path);
int cvt_unet_synchronize(cvt_unet *net);

/* Tuning: moment tile side (32, 64, 128 or 256; 0 = choose from the site
 * density).  Results do not depend on it (for sides >= 64 they are identical
 * bit for bit).  A tile can meet at most 1024, 256, 512 or 1024 distinct
 * sites respectively; more is reported as CVT_ERR_CAPACITY. */
int cvt_set_tile(cvt_ctx *ctx, int tile);
int cvt_get_tile(cvt_ctx *ctx);
/* Launch shape of the JFA passes, a speed choice (the labels do not depend
 * on it).  runs_mask < 0: runs of four rows one step apart where every wave
 * of the launch gets one, bands of rows elsewhere (the default); otherwise
 * bit k set = runs at step 2^k, clear = bands.  blocks_per_cu: workgroups
 * launched per CU (1 .. 16), 0 = the built-in grid (four 256-thread
 * workgroups per CU while the site table takes at most 40 KB of LDS, one
 * 1024-thread workgroup per CU above).  Replaces the fixed per-size launch
 * tables of pba2DInitializeInput (source/gcvt.cu:842-847); set by the
 * autotuner.  Takes effect at the next cvt_voronoi / cvt_lloyd call.
 * cvt_get_jfa_shape reports the steps that take runs under the current
 * shape (bit k for step 2^k; 0 where W is not a multiple of 4, which has a
 * per-pixel layout of its own) and the blocks_per_cu last set. */
int cvt_set_jfa_shape(cvt_ctx *ctx, int runs_mask, int blocks_per_cu);
int cvt_get_jfa_shape(cvt_ctx *ctx, int *runs_mask, int *blocks_per_cu);

#ifdef __cplusplus
}
#endif

#endif /* CVT_H */
