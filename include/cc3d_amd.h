/*
 * cc3d_amd.h -- C ABI of the MI355X labelling library (libcc3d_amd.so).
 *
 * Drop-in boundary for the labelling path of connected-components-3d 4.0.0.
 * Every pointer named `in`/`out` is a device (HBM) pointer; `stream` is a
 * hipStream_t (NULL = the legacy default stream).  Volumes are dense with x
 * fastest: index = x + sx * (y + sy * z), which is how the reference's
 * binding hands buffers to its C++ core (cc3d/fastcc3d.pyx:352-359).
 *
 * dtype codes: 0 u8, 1 u16, 2 u32, 3 u64, 4 f32, 5 f64.  Signed integers,
 * bool and float16 (delta == 0) are passed as the unsigned view of the same
 * width, exactly as the reference binding does (cc3d/fastcc3d.pyx:471-613).
 * Output dtypes: 1 u16, 2 u32, 3 u64.
 *
 * Status codes: 0 ok; 1 runtime error (the reference's std::runtime_error,
 * e.g. "maximum length exception"); 2 invalid argument; 3 HIP error;
 * 4 out of memory.  cc3d_amd_last_error() returns the calling thread's last
 * message.
 */
#ifndef CC3D_AMD_H
#define CC3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CC3D_AMD_OK 0
#define CC3D_AMD_ERR_RUNTIME 1
#define CC3D_AMD_ERR_INVALID 2
#define CC3D_AMD_ERR_HIP 3
#define CC3D_AMD_ERR_NOMEM 4

/* Replaces cc3d::connected_components3d<T,U>(T* in, sx, sy, sz, max_labels,
 * connectivity, T delta, U* out, size_t& N, bool periodic_boundary,
 * bool binary_image) -- cc3d/cc3d_continuous.hpp:394-455, bound by the
 * reference at cc3d/fastcc3d.pyx:67-74.  `delta` points to one host value of
 * the input type.  Every output voxel is written (the reference relies on a
 * pre-zeroed buffer).  max_labels sizes the reference's union-find; the same
 * overflow ("maximum length exception") is reported here as status 1. */
int cc3d_amd_connected_components3d(const void* in, int in_dtype, int64_t sx, int64_t sy,
                                    int64_t sz, int64_t max_labels, int64_t connectivity,
                                    const void* delta, void* out, int out_dtype, uint64_t* N,
                                    int periodic_boundary, int binary_image, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cc3d_amd_connected_components3d with the same arguments
 * and results, but every workspace byte lives in a fixed device array of
 * CC3D_AMD_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), and
 * the call makes no hipMalloc/hipFree or other allocating call.  Volumes
 * whose workspace (cc3d_amd_workspace_bytes) exceeds it are rejected with
 * status 2, e.g. 256^3 voxels fit.  All calls on one GPU share that array,
 * so calls that overlap in time on one GPU race: serialise them.  The call
 * returns after its work on `stream` has finished. */
#define CC3D_AMD_NOALLOC_SCRATCH_BYTES (256ull << 20)
int cc3d_amd_connected_components3d_noalloc(const void* in, int in_dtype, int64_t sx, int64_t sy,
                                            int64_t sz, int64_t max_labels, int64_t connectivity,
                                            const void* delta, void* out, int out_dtype, uint64_t* N,
                                            int periodic_boundary, int binary_image, void* stream);

/* Replaces cc3d::estimate_provisional_label_count<T>(T* in, sx, voxels,
 * int64& first_foreground_row, int64& last_foreground_row) --
 * cc3d/cc3d.hpp:287-315, bound at cc3d/fastcc3d.pyx:60-65. */
int cc3d_amd_estimate_provisional_label_count(const void* in, int in_dtype, int64_t sx,
                                              int64_t voxels, uint64_t* epl,
                                              int64_t* first_foreground_row,
                                              int64_t* last_foreground_row, void* stream);

/* Replaces epl_special_row (cc3d/fastcc3d.pyx:644-679), the binding's
 * shortcut for a volume whose only foreground is row `row` (y + sy * z) when
 * delta == 0 (cc3d/fastcc3d.pyx:469-470).  Writes every voxel of `out`: the
 * row's labels, 0 elsewhere.  The reference keeps the run value in a C
 * int64, so a voxel continues a run when it equals the truncation of the
 * previous voxel; *N receives the number of labels.  *fail_kind reports the
 * first start voxel whose int64 conversion raises in the reference (0 none,
 * 1 NaN -> ValueError, 2 infinity -> OverflowError, 3 out of int64 range ->
 * OverflowError; uint64 inputs are checked when check_u64 != 0) and
 * *fail_rank that start's label.  The caller raises: the conversion error
 * when fail_rank fits out_dtype, else numpy's out-of-bounds OverflowError
 * when N does not fit.  in_dtype as above (signed inputs as their unsigned
 * view). */
int cc3d_amd_special_row(const void* in, int in_dtype, int check_u64, int64_t sx, int64_t voxels, int64_t row,
                         void* out, int out_dtype, uint64_t* N, int* fail_kind, uint64_t* fail_rank,
                         void* stream);
const char* cc3d_amd_special_row_last_error(void);

/* ---- two-stage form used by the Python binding ----------------------------
 * The binding layer (cc3d/fastcc3d.pyx:244-626) needs the provisional-label
 * estimate before it picks the output dtype.  Stage 1 labels the volume and
 * computes that estimate in the same pass; stage 2 writes labels in the
 * chosen dtype.  No reference counterpart: these split the one reference call
 * so the estimate costs no extra pass over the input. */
typedef struct cc3d_amd_job cc3d_amd_job;

typedef struct {
  uint64_t N;                  /* number of components */
  uint64_t provisional;        /* provisional labels the reference would create */
  uint64_t epl;                /* estimate_provisional_label_count */
  int64_t first_foreground_row;
  int64_t last_foreground_row;
} cc3d_amd_stats;

/* flags for the two-stage form */
#define CC3D_AMD_CONTINUOUS 1          /* |a-b| <= delta linking (the delta != 0 path) */
#define CC3D_AMD_COUNT_PROVISIONAL 2   /* binary 2D-8: count provisional block labels */
#define CC3D_AMD_NO_SINGLE_ROW_RULE 4  /* do not mimic the binding's single-row shortcut */
/* Tile family of the local labelling kernel for 3D volumes (no reference
 * counterpart: a performance choice, results are identical): row tiles
 * (64x8x8, x-runs as union-find nodes), voxel tiles (32x8x8, one node per
 * voxel) or tall voxel tiles (32x16x8, one node per voxel, 512 threads).
 * None: row tiles for 6-connectivity, voxel tiles otherwise. */
#define CC3D_AMD_TILES_ROWS 8
#define CC3D_AMD_TILES_VOXELS 16
#define CC3D_AMD_TILES_TALL 32

/* Bytes of scratch a job of this geometry needs (0 on bad arguments). */
size_t cc3d_amd_workspace_bytes(int64_t sx, int64_t sy, int64_t sz, int64_t connectivity,
                                int binary_image, int flags);

/* Stage 1.  workspace may be NULL (the library allocates it stream-ordered);
 * a caller workspace must be 256-byte aligned.
 * With periodic_boundary on a multilabel volume whose foreground is a single
 * row, the wrap-around unions are skipped, as the binding's single-row
 * shortcut (cc3d/fastcc3d.pyx:469-470, :644-679) ignores periodicity, unless
 * CC3D_AMD_NO_SINGLE_ROW_RULE is set. */
int cc3d_amd_label_begin(cc3d_amd_job** job, const void* in, int in_dtype, int64_t sx,
                         int64_t sy, int64_t sz, int64_t connectivity, const void* delta,
                         int flags, int periodic_boundary, int binary_image,
                         void* workspace, size_t workspace_bytes, void* stream);
/* Waits for stage 1 and returns its counters. */
int cc3d_amd_job_stats(cc3d_amd_job* job, cc3d_amd_stats* stats);
/* Stage 2: writes every voxel of `out` (out_dtype 1/2/3). */
int cc3d_amd_label_finish(cc3d_amd_job* job, void* out, int out_dtype);
/* Releases the job (and its workspace when the library allocated it). */
void cc3d_amd_job_free(cc3d_amd_job* job);

/* Union-find length the reference's first pass would allocate for these
 * arguments (cc3d/cc3d.hpp:360-366 and twins, cc3d/cc3d_binary.hpp:47-51,
 * :350-352, :527-530, :794-797, :964-967); 0 when that pass returns before
 * labelling (max_labels == 0).  A provisional count >= this length is the
 * reference's "maximum length exception" (cc3d/cc3d.hpp:107-111). */
uint64_t cc3d_amd_union_find_length(int64_t max_labels, int64_t voxels, int64_t connectivity,
                                    int binary_image, int out_dtype);

/* ---- batches: many volumes labelled by one set of launches ---------------
 * No reference counterpart: the reference labels one volume per call.  Each
 * volume's counters and labels equal cc3d_amd_label_begin/_finish's on that
 * volume alone (default flags).  Every launch (local labelling, cross-tile
 * unions, root scan, finalize) covers all tiles of all volumes, so small
 * volumes fill the GPU and a batch costs one host synchronisation.  One mode
 * per batch: multilabel, binary or continuous (one delta for all volumes).
 * Per volume: connectivity 26, 18 or 6, or 8 and 4 on a single plane, and
 * optionally a periodic boundary (not continuous).  The sum of the volumes'
 * keys (voxels; 4 per 2x2 block for binary 2D-8) must stay below
 * 2^32 - 32 * n. */
typedef struct {
  const void* in;          /* device pointer, dense, x fastest */
  int64_t sx, sy, sz;
  int64_t connectivity;    /* | (CC3D_AMD_TILES_ROWS, _VOXELS or _TALL) << 8 to choose the tile family,
                              | CC3D_AMD_VOLUME_PERIODIC for a periodic boundary */
} cc3d_amd_volume;

#define CC3D_AMD_VOLUME_PERIODIC (1 << 16)
/* mode of a batch */
#define CC3D_AMD_BATCH_MULTILABEL 0
#define CC3D_AMD_BATCH_BINARY 1
#define CC3D_AMD_BATCH_CONTINUOUS 2

typedef struct cc3d_amd_batch cc3d_amd_batch;

/* Bytes of scratch a batch needs (0 if a volume is unsupported). */
size_t cc3d_amd_batch_workspace_bytes(const cc3d_amd_volume* vols, int n, int mode);
/* Plans a batch (no launches).  workspace may be NULL (allocated here and
 * freed by cc3d_amd_batch_free); a caller workspace must be 256-byte aligned.
 * All later launches go to `stream`. */
int cc3d_amd_batch_create(cc3d_amd_batch** batch, const cc3d_amd_volume* vols, int n, int in_dtype,
                          int mode, void* workspace, size_t workspace_bytes, void* stream);
/* Continuous batches: delta (a value of in_dtype), before the first begin or
 * graph capture. */
int cc3d_amd_batch_set_delta(cc3d_amd_batch* batch, const void* delta);
/* Output buffer and dtype (1/2/3) of every volume; uploaded stream-ordered. */
int cc3d_amd_batch_set_outputs(cc3d_amd_batch* batch, void* const* outs, const int* out_dtypes);
/* Stage 1 for all volumes (labels, counters); no host synchronisation. */
int cc3d_amd_batch_begin(cc3d_amd_batch* batch);
/* Waits for the stream and returns n stats records. */
int cc3d_amd_batch_stats(cc3d_amd_batch* batch, cc3d_amd_stats* stats);
/* Stage 2 for all volumes into the buffers set by cc3d_amd_batch_set_outputs. */
int cc3d_amd_batch_finish(cc3d_amd_batch* batch);
/* Captures begin + finish into a HIP graph (needs a caller workspace and
 * outputs already set); cc3d_amd_batch_graph_launch replays it on the stream.
 * Replays read the current contents of the inputs and the output table. */
int cc3d_amd_batch_graph_capture(cc3d_amd_batch* batch);
int cc3d_amd_batch_graph_launch(cc3d_amd_batch* batch);
void cc3d_amd_batch_free(cc3d_amd_batch* batch);

/* ---- follow-ons: statistics and dust ------------------------------------ */

/* Max (unsigned view) or signed min/max of a label volume in HBM; used for
 * N = np.max(labels) and the sign check of cc3d/fastcc3d.pyx:731-748. */
int cc3d_amd_label_range(const void* labels, int dtype, int64_t voxels, int is_signed,
                         uint64_t* umax, int64_t* smin, int64_t* smax, void* stream);

/* Replaces cc3d.statistics (cc3d/fastcc3d.pyx:681-938).  Axes are the numpy
 * axes (s0, s1, s2; s2 = 1 for ndim == 2); f_order selects which axis is
 * fastest in memory.  Outputs (device): counts uint32[N+1];
 * bbox uint32[(N+1)*2*ndim] as min,max per axis (min 0xFFFFFFFF, max 0 for
 * absent labels); centroids float64[(N+1)*ndim] (NaN for absent labels);
 * sums_scratch uint64[(N+1)*ndim].  bbox and centroids may be NULL: that
 * output is then not computed (centroids need sums_scratch). */
int cc3d_amd_statistics(const void* labels, int dtype, int64_t s0, int64_t s1, int64_t s2,
                        int ndim, int f_order, uint64_t N, uint32_t* counts, uint32_t* bbox,
                        double* centroids, unsigned long long* sums_scratch, void* stream);

/* The masking step of cc3d.dust (cc3d/__init__.py:121-150): zero img where
 * the voxel's component fails `size < lo` (ranged == 0) or
 * `!(lo <= size < hi)` (ranged != 0), inverted when invert != 0.
 * erase_scratch: uint8[N+1]; *n_masked (device uint64, pre-zeroed) receives
 * the number of labels in 1..N that fail the test. */
int cc3d_amd_dust_apply(void* img, int img_dtype, const void* labels, int label_dtype,
                        int64_t voxels, const uint32_t* counts, uint64_t N, double lo, double hi,
                        int ranged, int invert, uint8_t* erase_scratch,
                        unsigned long long* n_masked, void* stream);
/* cc3d_amd_dust_apply into a separate image: dst = src with the failing
 * components zeroed, the copy and the masking of cc3d/__init__.py:121-150 in
 * one pass (dst and src must not overlap).  For callers that keep the
 * original image; the arguments otherwise mean the same. */
int cc3d_amd_dust_copy(const void* src, void* dst, int img_dtype, const void* labels, int label_dtype,
                       int64_t voxels, const uint32_t* counts, uint64_t N, double lo, double hi,
                       int ranged, int invert, uint8_t* erase_scratch,
                       unsigned long long* n_masked, void* stream);

/* The renumbering step of cc3d.largest_k without fastremap
 * (cc3d/__init__.py:250-275): out[v] = i + 1 where labels[v] == keep[i], else
 * 0 (uint8 output with k == 1: the boolean mask).  keep: device uint64[k], the
 * labels to keep in output order; lut_scratch: device uint32[N+1]; every label
 * is <= N. */
int cc3d_amd_keep_labels(const void* labels, int label_dtype, int64_t voxels, const uint64_t* keep,
                         int64_t k, uint64_t N, uint32_t* lut_scratch, void* out, int out_dtype,
                         void* stream);
const char* cc3d_amd_follow_last_error(void);

/* ---- voxel connectivity graph ------------------------------------------ */

/* Replaces extract_voxel_connectivity_graph<T, OUT> (cc3d/cc3d_graphs.hpp:
 * 192-236, bound at cc3d/fastcc3d.pyx:1021-1176).  labels: device, x fastest,
 * an unsigned view (dtype 0-3); out: device, (sx*sy*sz) uint32 for
 * connectivity 26/18, uint8 for 6/8/4.  Bit d is set unless the neighbour in
 * direction d exists and holds a different label; the bit order is the
 * reference's (cc3d/fastcc3d.pyx:1036-1064).  4/8 need sz == 1 (status 1,
 * "sz must be 1 for 2D connectivities."). */
int cc3d_amd_voxel_connectivity_graph(const void* labels, int dtype, int64_t sx, int64_t sy, int64_t sz,
                                      int connectivity, void* out, void* stream);
const char* cc3d_amd_graphs_last_error(void);

/* Replaces color_connectivity_graph_N<VCG_t, uint32_t>
 * (cc3d/cc3d_graphs.hpp:1077-1106, bound at cc3d/fastcc3d.pyx:941-1018):
 * connected components of a voxel connectivity graph.  vcg: device, x
 * fastest, uint8 (vcg_size 1) or uint32 (4) with the bit layout of
 * cc3d_amd_voxel_connectivity_graph; every voxel is labelled (1..N, numbered
 * by first voxel in raster order); voxel p joins its backward neighbour q
 * when p's bit for the direction p -> q is set.  connectivity 26 or 6 for
 * sz > 1 (26 needs a uint32 graph), 8 or 4 (6 and 26 read as 4 and 8) for
 * sz == 1.  out: device uint32 (sx*sy*sz).  Waits for the stream. */
int cc3d_amd_color_connectivity_graph(const void* vcg, int vcg_size, int64_t sx, int64_t sy, int64_t sz,
                                      int connectivity, uint32_t* out, uint64_t* N, void* stream);
const char* cc3d_amd_color_last_error(void);

/* Replaces extract_region_graph<T> (cc3d/cc3d_graphs.hpp:430-468, with
 * extract_region_graph_2d/_3d at :305-428; bound at cc3d/fastcc3d.pyx:1198-1252
 * as contacts / region_graph): the label pairs in contact and, per pair, the
 * float32 sum the reference accumulates (face areas from wx, wy, wz, or
 * neighbour counts when surface_area is 0), bit for bit.  labels: device, x
 * fastest; dtype 0 u8, 1 u16, 2 u32, 3 u64, 6 i8, 7 i16, 8 i32, 9 i64 (pairs
 * are ordered by the type's own comparison).  connectivity 6/18/26, or 4/8
 * with sz == 1 (status 1, "z thickness must be 1 for 2d region graph
 * extraction.").  On success *result holds n_pairs pairs on the device. */
typedef struct cc3d_amd_contacts cc3d_amd_contacts;
int cc3d_amd_contacts_run(const void* labels, int dtype, int64_t sx, int64_t sy, int64_t sz,
                          int connectivity, int surface_area, float wx, float wy, float wz,
                          cc3d_amd_contacts** result, uint64_t* n_pairs, void* stream);
/* Copies the pairs to host arrays of n_pairs entries: lo < hi as 64-bit
 * words (signed labels sign-extended), and the sums. */
int cc3d_amd_contacts_fetch(cc3d_amd_contacts* result, uint64_t* lo, uint64_t* hi, float* value);
void cc3d_amd_contacts_free(cc3d_amd_contacts* result);
const char* cc3d_amd_contacts_last_error(void);

/* Replaces extract_runs<T> (cc3d/cc3d_graphs.hpp:470-501, bound at
 * cc3d/fastcc3d.pyx:1254-1277 as runs): the maximal runs of equal values of
 * the flat array labels (device, label_size 1/2/4/8 bytes, unsigned), as
 * (label, start, end) sorted by label, each label's runs in position order.
 * Label-0 runs are included (first); the reference leaves them out except
 * for a one-voxel array.  n < 2^32 - 1. */
typedef struct cc3d_amd_runs cc3d_amd_runs;
int cc3d_amd_extract_runs(const void* labels, int label_size, uint64_t n, cc3d_amd_runs** result, uint64_t* n_runs,
                          void* stream);
/* Device pointers to the n_runs (label, start, end) entries, valid until free. */
int cc3d_amd_runs_device(cc3d_amd_runs* result, const uint64_t** label, const uint64_t** start,
                         const uint64_t** end);
int cc3d_amd_runs_fetch(cc3d_amd_runs* result, uint64_t* label, uint64_t* start, uint64_t* end);
void cc3d_amd_runs_free(cc3d_amd_runs* result);
/* Replaces set_run_voxels<T> (cc3d/cc3d_graphs.hpp:503-523, bound at
 * cc3d/fastcc3d.pyx:1279-1308 as draw): image[start[k] .. end[k]) = value for
 * every run k (image: device, image_size 1/2/4/8 bytes, n elements; start,
 * end: device).  A run with end > n or start >= end stops the painting there:
 * the runs before it are painted and status 1 ("Invalid run.") returned. */
int cc3d_amd_draw(void* image, int image_size, uint64_t n, uint64_t value, const uint64_t* start,
                  const uint64_t* end, uint64_t n_runs, void* stream);
const char* cc3d_amd_runs_last_error(void);

/* Streaming a volume in z-slabs (cc3d/__init__.py:353-501,
 * connected_components_stack; no single-call counterpart: the reference
 * labels each slab with connected_components and merges seams in Python).
 * Slab k's labels 1..N_k become provisional ids offset_k + label; a 64-bit
 * forest over the ids (entry p = parent, union by the smaller id) joins the
 * seams; the final label of id p is 1 + the rank of its root.
 *
 * stack_seam: unions across the seam between slab k-1's last plane
 * (prev_values, prev_labels; device, sx*sy, x fastest) and slab k's first
 * plane, for 6/18/26-connectivity; labels are the slabs' own (uint32), the
 * offsets turn them into ids.  dtype as the labelling call (0 u8 .. 5 f64). */
int cc3d_amd_stack_seam(const void* prev_values, const uint32_t* prev_labels, const void* values,
                        const uint32_t* labels, int dtype, int64_t sx, int64_t sy, uint64_t prev_offset,
                        uint64_t offset, int connectivity, int binary_image, uint64_t* forest, void* stream);
/* Adds to *count (device) the foreground voxels of the first plane that take
 * a new provisional label in their slab's own pass but not in a pass over
 * the whole volume (they link to the previous slab's last plane). */
int cc3d_amd_stack_seam_provisional(const void* prev_values, const void* values, int dtype, int64_t sx,
                                    int64_t sy, int connectivity, int binary_image, uint64_t* count,
                                    void* stream);
/* forest[p] = p for first <= p < last (device). */
int cc3d_amd_forest_init(uint64_t* forest, uint64_t first, uint64_t last, void* stream);
/* lut[p] (device, n_ids + 1 entries) = final label of id p (lut[0] = 0);
 * *N = number of components (host).  Waits for the stream. */
int cc3d_amd_forest_labels(const uint64_t* forest, uint64_t n_ids, uint64_t* lut, uint64_t* N, void* stream);
/* out[i] = labels[i] ? labels[i] + offset : 0 (uint32 in, uint64 out, device). */
int cc3d_amd_offset_labels(const uint32_t* labels, uint64_t n, uint64_t offset, uint64_t* out, void* stream);
/* out[i] = lut[ids[i]] in the output dtype (1 u16, 2 u32, 3 u64; device). */
int cc3d_amd_map_labels(const uint64_t* ids, uint64_t n, const uint64_t* lut, void* out, int out_dtype,
                        void* stream);
const char* cc3d_amd_stack_last_error(void);

/* Per-kernel timing with HIP events recorded on the launch stream.  No
 * reference counterpart (measurement support).  on: 0 off, 1 every kernel,
 * 2 only the local labelling kernel.  collect() waits for the events and
 * returns rows (name, total ms, launches) aggregated by kernel. */
void cc3d_amd_profile_enable(int on);
int cc3d_amd_profile_collect(char* names, int name_len, double* total_ms, int* counts, int cap);

const char* cc3d_amd_last_error(void);
const char* cc3d_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CC3D_AMD_H */
