/*
 * sat.h - C ABI of the MI355X (gfx950) summed-area-table library.
 *
 * Every entry point takes plain device pointers, element pitches and sizes,
 * returns a sat_status, never aborts and never frees caller memory. Streams
 * are passed as `void*` (a hipStream_t; NULL selects the default stream).
 * All work is stream-ordered; nothing here synchronises the host unless its
 * comment says so.
 *
 * Reference interfaces replaced (pengdada/SAT, paths relative to its root):
 *   BlockScan::blockScan       integralImage/ScanRow-BRLT.cu:24-25
 *   SerielScan::serielScan     integralImage/BRLT-Scan.cu:14-15
 *   IncrementScan::ScanX       integralImage/ScanRowColumn.cu:144-145
 *   IncrementScan::ScanY       integralImage/ScanRowColumn.cu:15-16
 *   <NS>::Test<TSrc,TDst,SCAN> two-pass full SAT compositions
 *                              ScanRow-BRLT.cu:225-226, BRLT-Scan.cu:237-238,
 *                              ScanRowColumn.cu:475-476
 *   DataType / SCAN enums      integralImage/integralImage.h:22-37
 *   algorithm ids              integralImage/main.cpp:20-34
 */
#ifndef SAT_AMD_SAT_H
#define SAT_AMD_SAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAT_ABI_VERSION 2

typedef enum sat_status {
  SAT_OK = 0,
  SAT_ERR_NULL = 1,          /* a required pointer is NULL                         */
  SAT_ERR_INVALID = 2,       /* bad size, pitch, count, box or algorithm id        */
  SAT_ERR_DTYPE = 3,         /* dtype pair not supported by this entry point       */
  SAT_ERR_WORKSPACE = 4,     /* caller workspace too small                         */
  SAT_ERR_HIP = 5,           /* a HIP runtime call failed (launch, alloc, copy)    */
  SAT_ERR_UNSUPPORTED = 6,   /* valid request this build does not implement        */
  SAT_ERR_LOOKBACK = 7       /* a look-back gave up waiting for a carry (see below) */
} sat_status;

/* Numerically equal to the reference's DataType (integralImage.h:22-32). */
typedef enum sat_dtype {
  SAT_U8 = 0,
  SAT_I8 = 1,
  SAT_U16 = 2,
  SAT_I16 = 3,
  SAT_U32 = 4,
  SAT_I32 = 5,
  SAT_F32 = 6,
  SAT_F64 = 7
} sat_dtype;

/* Numerically equal to the reference's SCAN (integralImage.h:34-37). Selects
 * the in-wave scan order (Kogge-Stone or Ladner-Fischer); integer results do
 * not depend on it, floating-point results may differ in rounding. */
typedef enum sat_scan { SAT_SCAN_KOGGE_STONE = 0, SAT_SCAN_LF = 1 } sat_scan;

/* 0..2 are the reference's algorithm ids (main.cpp:20-34); each composes two
 * passes through a temporary like the reference's Test<> harnesses. 3 is this
 * library's one-pass kernel (reads the input once, writes the output once). */
typedef enum sat_alg {
  SAT_ALG_SCANROW_BRLT = 0,   /* BlockScan:     blockScan(blockScan(A))        */
  SAT_ALG_BRLT_SCANROW = 1,   /* SerielScan:    serielScan(serielScan(A))      */
  SAT_ALG_SCANROW_COLUMN = 2, /* IncrementScan: ScanY(ScanX(A))                */
  SAT_ALG_SINGLE_PASS = 3     /* one launch, decoupled look-back in 2-D        */
} sat_alg;

/* Supported (TSrc, TDst) pairs: the reference dispatchers' set
 * (ScanRow-BRLT.cu:274-376): u8->{u32,f32,f64}, i8->{i32,f32,f64},
 * i32->{i32,f64}, f32->{f32,f64}, f64->f64; plus u32->u32. Integer outputs
 * wrap modulo 2^32. */
int sat_dtype_pair_supported(int tin, int tout);
const char* sat_status_string(int status);
int sat_abi_version(void);
/* Size in bytes of one element of `dtype`, 0 if unknown. */
size_t sat_dtype_size(int dtype);

/* Look-back give-ups. A launch never spins forever: a look-back whose carry
 * does not arrive within about a second gives up and the launch finishes;
 * the results of that launch are then undefined. A give-up is recorded in
 * two sticky places, both kept until they are read here:
 *   - the fault word of the launch's workspace;
 *   - a flag per device in pinned host memory.
 * The drop-ins never fail a later call because of an earlier give-up; callers
 * check with one of these (the Python wrappers check the device flag at every
 * call and raise).
 *
 * sat_workspace_fault: reports and clears the fault word of a workspace
 * (NULL: the library-managed carry workspaces of the calling device and
 * `stream`): 1 if some look-back on it gave up since the last query, 0 if
 * none did, a negative sat_status on failure. Synchronises `stream`. */
int sat_workspace_fault(void* workspace, void* stream);

/* Reads the calling device's host-visible give-up flag without
 * synchronising: 1 if a look-back of any launch of this library on the
 * device that has finished gave up since the flag was last cleared, else 0;
 * clears it when `clear` is nonzero. Negative sat_status on failure. */
int sat_device_fault(int clear);

/* Frees every library-managed workspace of the calling device after
 * synchronising it. Managed workspaces are otherwise kept until the process
 * ends: a buffer replaced by a larger one is retired, not freed, so that
 * graphs captured with it and launches still in flight keep valid memory.
 * Call this only when no captured graph that used the managed workspaces
 * will be replayed again. */
int sat_release_workspaces(void);

/* Streaming device-to-device copy of `bytes` bytes (16-byte aligned,
 * non-overlapping buffers), stream-ordered. No counterpart in the reference:
 * it is the measured bandwidth ceiling the SAT kernels are compared against
 * (the same 16-byte vector loads and nontemporal stores, no carries). */
int sat_stream_copy(const void* src, void* dst, size_t bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Pass-level drop-ins. Pitches (`widthStride`, `heightStride`) are in elements
 * of the pointed-to type, as GpuData::DataPitch() returns them
 * (cudaLib.cuh:178-180). The tuning template parameters of the reference
 * kernels (BLOCK_SIZE, SMEM_COUNT, BLOCK_DIM_*, COUNT_PER_THREAD, WARP_COUNT)
 * have no counterpart: tiles are chosen for wave64 inside the library. These
 * calls use a library-managed carry workspace per (device, stream), which is
 * allocated (hipMalloc, once, growing) on first use.
 * ------------------------------------------------------------------------- */

/* BlockScan::blockScan (ScanRow-BRLT.cu:24-25): row scan, then
 * block-register-local transpose; out[x*heightStride + y] = sum_{k<=x} in[y*widthStride + k].
 * `in` is height x width, `out` is width x height. */
int sat_block_scan(int scan, int tin, int tout, const void* in, void* out, uint32_t width,
                   uint32_t widthStride, uint32_t height, uint32_t heightStride, void* stream);

/* SerielScan::serielScan (BRLT-Scan.cu:14-15): block-register-local
 * transpose, then serial row scan; same result layout as sat_block_scan. The
 * reference kernel has no scan-type parameter (BRLT-Scan.cu:14). */
int sat_seriel_scan(int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
                    uint32_t height, uint32_t heightStride, void* stream);

/* IncrementScan::ScanX (ScanRowColumn.cu:144-145): row scan in place layout,
 * out[y*widthStride + x] = sum_{k<=x} in[y*widthStride + k]; one stride for
 * both buffers, counted in elements of each buffer's own type. */
int sat_scan_x(int scan, int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
               uint32_t height, void* stream);

/* IncrementScan::ScanY (ScanRowColumn.cu:15-16): column scan,
 * out[y*widthStride + x] = sum_{k<=y} in[k*widthStride + x]. */
int sat_scan_y(int scan, int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
               uint32_t height, void* stream);

/* Row / column scans with separate input and output pitches (elements). */
int sat_scan_rows(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                  uint32_t width, uint32_t height, void* stream);
int sat_scan_cols(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                  uint32_t width, uint32_t height, void* stream);

/* ---------------------------------------------------------------------------
 * Full summed-area table: out[y][x] = sum_{k<=y, l<=x} in[k][l]
 * (IntegralImageSerial, integralImage.h:105-124). `in` and `out` are
 * height x width, pitches in elements of their own types. `workspace` may be
 * NULL (library-managed) or caller device memory of at least
 * sat_integral_workspace_bytes() bytes, 16-byte aligned (SAT_ERR_INVALID
 * otherwise; hipMalloc memory is), zero-filled once before its first use and
 * not shared by calls that can overlap. The same holds for the workspaces of
 * the reverse SAT, the batch and sat_box_grad.
 * ------------------------------------------------------------------------- */
size_t sat_integral_workspace_bytes(int alg, int tin, int tout, uint32_t width, uint32_t height);
int sat_integral(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out,
                 int64_t out_pitch, uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes,
                 void* stream);

/* Same over `planes` images of one shape, plane strides in elements (e.g. the
 * N*C planes of an NCHW tensor). */
size_t sat_integral_planes_workspace_bytes(int alg, int tin, int tout, uint32_t width, uint32_t height,
                                           uint32_t planes);
int sat_integral_planes(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch,
                        int64_t in_plane_stride, void* out, int64_t out_pitch, int64_t out_plane_stride,
                        uint32_t width, uint32_t height, uint32_t planes, void* workspace, size_t workspace_bytes,
                        void* stream);

/* One-pass SAT (algorithm 3) that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device or pinned memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Its carry workspace is a device array of
 * sat_integral_noalloc_max_bytes() (32 MiB) bytes that the library holds on
 * every device it runs on until the process ends; shapes whose workspace
 * (sat_integral_workspace_bytes(3, ...)) exceeds it are refused with
 * SAT_ERR_WORKSPACE (about 200 million pixels fit, f32 14000 x 14000). Calls on one
 * device share that array: calls that can overlap on one device (other
 * streams, other host threads) race and must not be issued. Arguments and
 * results as sat_integral with alg = 3. */
size_t sat_integral_noalloc_max_bytes(void);
int sat_integral_noalloc(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                         uint32_t width, uint32_t height, void* stream);

/* OpenCV-style layout (cv::integral): `out` is (height+1) x (width+1) with
 * row 0 and column 0 zero and out[y+1][x+1] = sum_{k<=y, l<=x} in[k][l];
 * out_pitch >= width + 1 elements. Same algorithms, pairs and workspace as
 * sat_integral (sat_integral_workspace_bytes of width x height). The
 * reference's README compares against OpenCV's integral (README.md:29); no
 * reference entry point has this layout. */
int sat_integral_bordered(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out,
                          int64_t out_pitch, uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes,
                          void* stream);
int sat_integral_bordered_planes(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch,
                                 int64_t in_plane_stride, void* out, int64_t out_pitch, int64_t out_plane_stride,
                                 uint32_t width, uint32_t height, uint32_t planes, void* workspace,
                                 size_t workspace_bytes, void* stream);

/* Reverse-direction SAT (the adjoint of sat_integral):
 * out[y][x] = sum_{k>=y, l>=x} in[k][l]. One pass. tin == tout, f32 or f64
 * (also u32 / i32, which wrap). No reference counterpart. */
size_t sat_integral_reverse_workspace_bytes(int dtype, uint32_t width, uint32_t height);
int sat_integral_reverse(int scan, int dtype, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                         uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes, void* stream);

size_t sat_integral_reverse_planes_workspace_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes);
int sat_integral_reverse_planes(int scan, int dtype, const void* in, int64_t in_pitch, int64_t in_plane_stride,
                                void* out, int64_t out_pitch, int64_t out_plane_stride, uint32_t width,
                                uint32_t height, uint32_t planes, void* workspace, size_t workspace_bytes,
                                void* stream);

/* ---------------------------------------------------------------------------
 * Batched SAT of many images of uneven sizes in ONE launch walking a queue
 * of (image, tile) items. No reference counterpart.
 * Images may mix dtype pairs whose outputs have the same width (all 4-byte
 * or all 8-byte outputs).
 * ------------------------------------------------------------------------- */
typedef struct sat_image {
  const void* in;
  void* out;
  int64_t in_pitch;  /* elements of tin  */
  int64_t out_pitch; /* elements of tout */
  uint32_t width;
  uint32_t height;
  int32_t tin;
  int32_t tout;
} sat_image;

size_t sat_batch_workspace_bytes(const sat_image* images, int count);
/* Validates the batch, builds the work queue on the host and copies it into
 * the workspace (blocking copy; call outside graph capture). */
int sat_batch_prepare(const sat_image* images, int count, void* workspace, size_t workspace_bytes, void* stream);
/* Runs the prepared queue in a single launch (preceded by a one-workgroup
 * epoch bump). Allocates nothing, copies nothing: may be captured in a
 * hipGraph and replayed. */
int sat_batch_run(void* workspace, void* stream);
/* Form of the launch for plans prepared from now on: 0 (default) one
 * workgroup per queue item, which the hardware dispatches in queue order; 1
 * a persistent kernel, as many workgroups as are resident, each drawing the
 * next (image, tile) item of its XCD's queue from a ticket counter. Results
 * are the same; the default measured faster on MI355X. A batch whose whole
 * images would load the XCDs unevenly (its largest images are then laid over
 * all XCDs) always takes the first form. */
int sat_batch_persistent_set(int on);
int sat_batch_persistent_get(void);

/* ---------------------------------------------------------------------------
 * Tile configurations and the autotuner's table. No reference counterpart:
 * the reference fixes its block shapes through template parameters
 * (BLOCK_SIZE, SMEM_COUNT, BLOCK_DIM_X/Y, COUNT_PER_THREAD, WARP_COUNT,
 * e.g. ScanRow-BRLT.cu:24, ScanRowColumn.cu:16,144), which the drop-in
 * entry points accept but do not use. Here the one-pass SAT (forward and
 * reverse) and the batch kernel pick a wave64 tile per launch (tile shapes
 * for 4-byte outputs; 8-byte outputs take half the rows):
 *   0: 4 waves x 16 rows, 64 x 256 tiles, 4 waves per SIMD
 *   1: 8 waves x 16 rows, 128 x 256 tiles, 4 waves per SIMD
 *   2: 8 waves x 8 rows, 64 x 256 tiles, up to 6 waves per SIMD
 *   3: 8 waves x 8 rows, 64 x 256 tiles, 8 waves per SIMD (64 registers)
 *   4: 4 waves x 24 rows, 96 x 256 tiles, 3 waves per SIMD (168 registers)
 *   5: 1-byte inputs into 4-byte outputs: 4 waves x 64 rows kept packed (4
 *      input bytes per register), 256 x 256 tiles; other pairs as 1
 * (the batch kernel takes configurations 0-4)
 * from a table indexed by input width (1, 4 or 8 bytes), output width (4 or
 * 8 bytes) and shape class
 *   class_x = min(floor(log2(ceil(width / 256))), 8)
 *   class_y = min(floor(log2(ceil(height / 64))), 11)
 *   class_p = 0, 1, 2, 3 for 1, 2-3, 4-15, 16+ planes.
 * The table starts as the tuned table shipped with the build (measured by
 * sat_amd.tune on the target device and compiled in). A class with no entry
 * for its input width takes the entry of its output width's same-width pair
 * (e.g. u8 -> f64 takes f64 -> f64's), then the built-in rule (narrow images
 * of >= 4096 rows take configuration 1, the rest 0); cfg = -1 unsets an
 * entry. Results of integer pairs are identical in every configuration.
 * Workspace sizes returned by the *_workspace_bytes functions fit every
 * configuration. in_bytes must not exceed out_bytes.
 * ------------------------------------------------------------------------- */
int sat_tile_config_count(void);
void sat_tune_class(uint32_t width, uint32_t height, uint32_t planes, int* class_x, int* class_y, int* class_p);
/* configuration a class uses now (negative sat_status on bad arguments) */
int sat_tune_get(int in_bytes, int out_bytes, int class_x, int class_y, int class_p);
/* cfg = -1 unsets the entry */
int sat_tune_set(int in_bytes, int out_bytes, int class_x, int class_y, int class_p, int cfg);
/* configuration a launch of this shape and these element widths would use */
int sat_tune_pick(int in_bytes, int out_bytes, uint32_t width, uint32_t height, uint32_t planes);
/* batch kernel configuration per output width (used by the next sat_batch_prepare) */
int sat_batch_tune_get(int out_bytes);
int sat_batch_tune_set(int out_bytes, int cfg);

/* ---------------------------------------------------------------------------
 * Box-filter readout on summed-area tables (f32 or f64), P planes.
 * boxes: P*nbox records of 4 int32 {y0, x0, y1, x1}, inclusive corners,
 * 0 <= y0 <= y1 < height, 0 <= x0 <= x1 < width.
 * out[p*nbox + b] = S(y1,x1) - S(y0-1,x1) - S(y1,x0-1) + S(y0-1,x0-1)
 * (terms with a -1 index are 0). No reference counterpart.
 * ------------------------------------------------------------------------- */
int sat_box_sums(int dtype, const void* sat, int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height,
                 uint32_t planes, const int32_t* boxes, uint32_t nbox, void* out, void* stream);

/* Adjoint of sat_box_sums with respect to the SAT: adds +g at (y1,x1) and
 * (y0-1,x0-1) and -g at (y0-1,x1) and (y1,x0-1) of each box (indices < 0
 * dropped) into grad_sat, which the caller initialises (usually to zero).
 * Coinciding corners accumulate with atomics, so f32 sums at a pixel hit by
 * several corners may differ in the last bits from run to run. */
int sat_box_scatter(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_sat,
                    int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                    void* stream);

/* Gradient of sum_b g[p*nbox+b] * boxsum_b(SAT(x_p)) with respect to x:
 * grad_x[p][y][x] = sum over boxes b of plane p containing (y, x) of g[p*nbox+b].
 * Computed as the reverse SAT of the boxes' corner impulses without
 * materialising the impulse image; the reverse SAT accumulates in f64 for
 * both dtypes and rounds once on store. grad_x is overwritten. Corner
 * impulses that land on one pixel are first summed in `dtype` with atomics:
 * where three or more coincide, f32 results may differ in the last bit from
 * run to run. */
size_t sat_box_grad_workspace_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes, uint32_t nbox);
int sat_box_grad(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_x, int64_t pitch,
                 int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes, void* workspace,
                 size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SAT_AMD_SAT_H */
