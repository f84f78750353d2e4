/*
 * sat_oracle.c - TEST INFRASTRUCTURE ONLY. CPU restatement of the reference's
 * integral-image routines (pengdada/SAT integralImage/integralImage.h), used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. Nothing in the sat_amd product loads this file.
 *
 * Each routine keeps the reference's fold order:
 *   oracle_serial_*    IntegralImageSerial   integralImage.h:105-124
 *   oracle_parallel_*  IntegralImageParallel integralImage.h:129-146 (OpenMP)
 *   oracle_allrow_*    IntegralImageAllRow   integralImage.h:148-160 (OpenMP)
 *   oracle_stride_*    IntegralStride        integralImage.h:90-100
 *   oracle_colscan_*   IntegralStride per column, stride = width
 *   oracle_transpose_* Transpose             integralImage.h:280-288
 * Accumulators have the output type and every pixel is converted to it
 * before the add (`sum += src[i]`, integralImage.h:79). Integer outputs add in
 * uint32_t so that overflow wraps modulo 2^32 (the reference's int
 * accumulators make signed overflow undefined; its uint ones wrap).
 *
 * Not in the reference (restated from the definitions):
 *   oracle_reverse_*   suffix SAT: out[y][x] = sum_{k>=y,l>=x} in[k][l]
 *   oracle_box_sums_*  four-corner box readout from a SAT
 *   sm64_*             splitmix64 counter-based input generators
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint8_t u8;
typedef int8_t i8;
typedef uint32_t u32;
typedef int32_t i32;
typedef float f32;
typedef double f64;

/* arithmetic type of each output type */
typedef u32 acc_u32;
typedef u32 acc_i32;
typedef f32 acc_f32;
typedef f64 acc_f64;

#define CVT(TD, x) ((acc_##TD)(TD)(x))

#define DEFINE_PAIR(TS, TD)                                                                       \
  /* IntegralRow (integralImage.h:74-82) */                                                       \
  static void row_##TS##_##TD(const TS* src, TD* dst, int64_t count) {                             \
    acc_##TD sum = 0;                                                                             \
    for (int64_t i = 0; i < count; i++) {                                                         \
      sum += CVT(TD, src[i]);                                                                     \
      dst[i] = (TD)sum;                                                                           \
    }                                                                                             \
  }                                                                                               \
  /* IntegralStride (integralImage.h:90-100) */                                                   \
  void oracle_stride_##TS##_##TD(const TS* src, int64_t stride_src, TD* dst, int64_t stride_dst,  \
                                 int64_t count) {                                                 \
    acc_##TD sum = 0;                                                                             \
    const TS* p0 = src;                                                                           \
    TD* p1 = dst;                                                                                 \
    for (int64_t i = 0; i < count; i++, p0 += stride_src, p1 += stride_dst) {                     \
      sum += CVT(TD, *p0);                                                                        \
      *p1 = (TD)sum;                                                                              \
    }                                                                                             \
  }                                                                                               \
  /* IntegralImageSerial (integralImage.h:105-124): row 0 is a left fold; */                      \
  /* every later row adds its own left fold to the row above. */                                  \
  void oracle_serial_##TS##_##TD(const TS* src, TD* dst, int64_t width, int64_t height) {          \
    acc_##TD sum = 0;                                                                             \
    const TS* p0 = src;                                                                           \
    TD* p1 = dst;                                                                                 \
    TD* p2 = dst;                                                                                 \
    for (int64_t j = 0; j < width; j++, p0++, p1++) {                                             \
      sum += CVT(TD, *p0);                                                                        \
      *p1 = (TD)sum;                                                                              \
    }                                                                                             \
    for (int64_t i = 1; i < height; i++) {                                                        \
      sum = 0;                                                                                    \
      for (int64_t j = 0; j < width; j++, p0++, p1++, p2++) {                                     \
        sum += CVT(TD, *p0);                                                                      \
        *p1 = (TD)((acc_##TD)(*p2) + sum);                                                        \
      }                                                                                           \
    }                                                                                             \
  }                                                                                               \
  /* IntegralImageParallel (integralImage.h:129-146): rows, then columns. */                      \
  void oracle_parallel_##TS##_##TD(const TS* src, TD* dst, int64_t width, int64_t height) {        \
    _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < height; i++)             \
        row_##TS##_##TD(src + i * width, dst + i * width, width);                                 \
    _Pragma("omp parallel for schedule(static)") for (int64_t j = 0; j < width; j++)              \
        oracle_stride_##TD##_##TD(dst + j, width, dst + j, width, height);                        \
  }                                                                                               \
  /* IntegralImageAllRow (integralImage.h:148-160) */                                             \
  void oracle_allrow_##TS##_##TD(const TS* src, TD* dst, int64_t width, int64_t height) {          \
    _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < height; i++)             \
        row_##TS##_##TD(src + i * width, dst + i * width, width);                                 \
  }                                                                                               \
  /* column-only scan: IntegralStride with stride = width on every column */                      \
  void oracle_colscan_##TS##_##TD(const TS* src, TD* dst, int64_t width, int64_t height) {         \
    _Pragma("omp parallel for schedule(static)") for (int64_t j = 0; j < width; j++)              \
        oracle_stride_##TS##_##TD(src + j, width, dst + j, width, height);                        \
  }

/* the same-type strides the parallel column pass needs come first */
DEFINE_PAIR(u32, u32)
DEFINE_PAIR(i32, i32)
DEFINE_PAIR(f32, f32)
DEFINE_PAIR(f64, f64)
DEFINE_PAIR(u8, u32)
DEFINE_PAIR(u8, f32)
DEFINE_PAIR(u8, f64)
DEFINE_PAIR(i8, i32)
DEFINE_PAIR(i8, f32)
DEFINE_PAIR(i8, f64)
DEFINE_PAIR(i32, f64)
DEFINE_PAIR(f32, f64)

/* Transpose (integralImage.h:280-288): dst is dst_height x dst_width with
 * dst_width == src_height. */
#define DEFINE_TRANSPOSE(T)                                                                         \
  void oracle_transpose_##T(const T* src, int64_t src_width, int64_t src_height, T* dst) {          \
    const int64_t dst_width = src_height, dst_height = src_width;                                  \
    for (int64_t y = 0; y < dst_height; y++)                                                        \
      for (int64_t x = 0; x < dst_width; x++) dst[x + y * dst_width] = src[y + x * src_width];      \
  }
DEFINE_TRANSPOSE(u32)
DEFINE_TRANSPOSE(i32)
DEFINE_TRANSPOSE(f32)
DEFINE_TRANSPOSE(f64)

/* Suffix SAT, the adjoint of the forward SAT: the mirror image of
 * IntegralImageSerial's fold (last row right-to-left, then each row above
 * adds its own right fold to the row below). */
#define DEFINE_REVERSE(T)                                                                           \
  void oracle_reverse_##T(const T* src, T* dst, int64_t width, int64_t height) {                    \
    for (int64_t i = height - 1; i >= 0; i--) {                                                     \
      acc_##T sum = 0;                                                                              \
      for (int64_t j = width - 1; j >= 0; j--) {                                                    \
        sum += (acc_##T)src[i * width + j];                                                         \
        dst[i * width + j] = (T)(i == height - 1 ? sum : (acc_##T)dst[(i + 1) * width + j] + sum);  \
      }                                                                                             \
    }                                                                                               \
  }
DEFINE_REVERSE(u32)
DEFINE_REVERSE(f32)
DEFINE_REVERSE(f64)

/* Box readout from a SAT plane set (f64 arithmetic): boxes are
 * {y0, x0, y1, x1} inclusive; out-of-image or empty boxes give NaN. */
void oracle_box_sums_f64(const f64* sat, int64_t width, int64_t height, int64_t planes, const i32* boxes,
                         int64_t nbox, f64* out) {
  for (int64_t p = 0; p < planes; p++) {
    const f64* s = sat + p * width * height;
    for (int64_t b = 0; b < nbox; b++) {
      const i32* q = boxes + 4 * (p * nbox + b);
      const int64_t y0 = q[0], x0 = q[1], y1 = q[2], x1 = q[3];
      if (y0 < 0 || x0 < 0 || y0 > y1 || x0 > x1 || y1 >= height || x1 >= width) {
        out[p * nbox + b] = NAN;
        continue;
      }
      const f64 a = s[y1 * width + x1];
      const f64 bb = y0 > 0 ? s[(y0 - 1) * width + x1] : 0.0;
      const f64 c = x0 > 0 ? s[y1 * width + x0 - 1] : 0.0;
      const f64 d = (y0 > 0 && x0 > 0) ? s[(y0 - 1) * width + x0 - 1] : 0.0;
      out[p * nbox + b] = ((a - bb) - c) + d;
    }
  }
}

/* ---------------------------------------------------------------------------
 * splitmix64 counter-based generators: r_i = mix(seed + (i+1) * golden).
 * ------------------------------------------------------------------------- */
static inline uint64_t sm64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t sm64_at(uint64_t seed, uint64_t i) { return sm64(seed, i); }

void sm64_u8(uint64_t seed, int64_t n, u8* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) out[i] = (u8)(sm64(seed, i) >> 56);
}
void sm64_i8(uint64_t seed, int64_t n, i8* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) out[i] = (i8)(u8)(sm64(seed, i) >> 56);
}
/* uniform in [0, 2^bits) */
void sm64_u32_bits(uint64_t seed, int64_t n, int bits, u32* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) out[i] =
      (u32)(sm64(seed, i) >> (64 - bits));
}
/* uniform in [0, 1) with 24 bits */
void sm64_f32_unit(uint64_t seed, int64_t n, f32* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) out[i] =
      (f32)((double)(sm64(seed, i) >> 40) * 0x1p-24);
}
/* normal(0,1) by Box-Muller on two 53-bit uniforms (counter 2i, 2i+1) */
void sm64_f64_normal(uint64_t seed, int64_t n, f64* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) {
    const double u1 = ((double)(sm64(seed, 2 * i) >> 11) + 1.0) * 0x1p-53; /* (0, 1] */
    const double u2 = (double)(sm64(seed, 2 * i + 1) >> 11) * 0x1p-53;     /* [0, 1) */
    out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
}
void sm64_f32_normal(uint64_t seed, int64_t n, f32* out) {
  _Pragma("omp parallel for schedule(static)") for (int64_t i = 0; i < n; i++) {
    const double u1 = ((double)(sm64(seed, 2 * i) >> 11) + 1.0) * 0x1p-53;
    const double u2 = (double)(sm64(seed, 2 * i + 1) >> 11) * 0x1p-53;
    out[i] = (f32)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
