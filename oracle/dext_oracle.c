/*
 * CPU oracle for the spot-finder hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this, and only as the checker. Compiled -O3 -ffp-contract=off so that no
 * multiply-add is fused: the reference is built for baseline x86-64, which has
 * no FMA, so its float tests round every product separately.
 *
 * Each function names the reference lines whose behaviour it restates.
 */
#include "dext_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Threshold: summed-area-table form, src/dext.cc:8-225.                     */
/*                                                                          */
/* T pixel, W wrapping accumulator, R real type (src/dext.hh:18-38).        */
/* Masked means value >= max(T) - 1 (src/dext.cc:40).                       */
/* ------------------------------------------------------------------------ */

#define KNL0 3 /* 7 x 7 dispersion window, src/dext.cc:37 */
#define ERODE 2 /* 5 x 5 erosion, src/dext.cc:117 */
#define KNL2 5 /* 11 x 11 Poisson window, src/dext.cc:165 */

/* Window sum over rows [i0+1, i1] x columns [j0+1, j1] of a summed-area
 * table, with i0 or j0 < 0 meaning "from the first row/column". Unsigned
 * arithmetic wraps, and the identity holds modulo 2^bits (src/dext.cc:79-95).
 */
#define SAT_WINDOW(W, sat, nx, i0, i1, j0, j1)                                 \
  ((W)((sat)[(size_t)(i1) * (nx) + (j1)] -                                    \
       ((i0) >= 0 ? (sat)[(size_t)(i0) * (nx) + (j1)] : (W)0) -               \
       ((j0) >= 0 ? (sat)[(size_t)(i1) * (nx) + (j0)] : (W)0) +               \
       (((i0) >= 0 && (j0) >= 0) ? (sat)[(size_t)(i0) * (nx) + (j0)] : (W)0)))

#define DEFINE_DEXT(SUFFIX, T, W, R, TMAX, SQRT)                               \
  int oracle_dext_##SUFFIX(const T *image, size_t height, size_t width,       \
                           oracle_signal_pixel *out, size_t capacity,         \
                           size_t *found) {                                   \
    *found = 0;                                                                \
    if (image == NULL || height == 0 || width == 0)                            \
      return -1;                                                               \
    if (height > (size_t)INT32_MAX / width)                                    \
      return -1;                                                               \
    const int32_t ny = (int32_t)height, nx = (int32_t)width;                   \
    const size_t n = height * width;                                           \
    const T masked = (T)(TMAX - 1);                                            \
    W *cnt = (W *)malloc(n * sizeof(W));                                       \
    W *sum = (W *)malloc(n * sizeof(W));                                       \
    W *sq = (W *)malloc(n * sizeof(W));                                        \
    uint8_t *mask0 = (uint8_t *)malloc(n);                                     \
    uint8_t *mask1 = (uint8_t *)malloc(n);                                     \
    if (!cnt || !sum || !sq || !mask0 || !mask1) {                             \
      free(cnt); free(sum); free(sq); free(mask0); free(mask1);                \
      return -4;                                                               \
    }                                                                          \
    /* Tables of valid count, value and value squared (src/dext.cc:53-65). */ \
    for (int32_t i = 0; i < ny; i++) {                                         \
      W rc = 0, rs = 0, rq = 0;                                                \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        const int ok = image[k] < masked;                                      \
        const W p = ok ? (W)image[k] : (W)0;                                   \
        rc += (W)ok;                                                           \
        rs += p;                                                               \
        rq += (W)(p * p);                                                      \
        cnt[k] = rc + (i ? cnt[k - nx] : (W)0);                                \
        sum[k] = rs + (i ? sum[k - nx] : (W)0);                                \
        sq[k] = rq + (i ? sq[k - nx] : (W)0);                                  \
      }                                                                        \
    }                                                                          \
    /* Dispersion test over the clipped 7 x 7 window (src/dext.cc:69-113). */ \
    for (int32_t i = 0; i < ny; i++) {                                         \
      const int32_t i0 = i - KNL0 - 1;                                         \
      const int32_t i1 = i + KNL0 < ny ? i + KNL0 : ny - 1;                    \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const int32_t j0 = j - KNL0 - 1;                                       \
        const int32_t j1 = j + KNL0 < nx ? j + KNL0 : nx - 1;                  \
        const W c = SAT_WINDOW(W, cnt, nx, i0, i1, j0, j1);                    \
        const W s = SAT_WINDOW(W, sum, nx, i0, i1, j0, j1);                    \
        const W q = SAT_WINDOW(W, sq, nx, i0, i1, j0, j1);                     \
        int hit = 0;                                                           \
        if (c >= 2) {                                                          \
          const R rm = (R)c, rs = (R)s, rq = (R)q;                             \
          const R lhs = rm * rq - rs * rs - rs * (rm - (R)1);                  \
          const R rhs = rs * (R)6 * SQRT((R)2 * (rm - (R)1));                  \
          hit = lhs > rhs;                                                     \
        }                                                                      \
        const size_t k = (size_t)i * nx + j;                                   \
        mask0[k] = (uint8_t)(hit || image[k] >= masked);                       \
      }                                                                        \
    }                                                                          \
    /* 5 x 5 erosion; off-frame neighbours count as set (src/dext.cc:117-142).\
     */                                                                        \
    for (int32_t i = 0; i < ny; i++) {                                         \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        uint8_t m = mask0[k];                                                  \
        if (image[k] >= masked) {                                              \
          m = 1;                                                               \
        } else if (m) {                                                        \
          for (int32_t a = i - ERODE; a <= i + ERODE && m; a++)                \
            for (int32_t b = j - ERODE; b <= j + ERODE && m; b++)              \
              if (a >= 0 && a < ny && b >= 0 && b < nx &&                      \
                  !mask0[(size_t)a * nx + b])                                  \
                m = 0;                                                         \
        }                                                                      \
        mask1[k] = m;                                                          \
      }                                                                        \
    }                                                                          \
    /* Tables of the pixels neither masked nor eroded-set (src/dext.cc:151-161)\
     */                                                                        \
    for (int32_t i = 0; i < ny; i++) {                                         \
      W rc = 0, rs = 0;                                                        \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        const int ok = image[k] < masked && !mask1[k];                         \
        rc += (W)ok;                                                           \
        rs += ok ? (W)image[k] : (W)0;                                         \
        cnt[k] = rc + (i ? cnt[k - nx] : (W)0);                                \
        sum[k] = rs + (i ? sum[k - nx] : (W)0);                                \
      }                                                                        \
    }                                                                          \
    /* Poisson test over the clipped 11 x 11 window at candidates, emitted in \
     * scan order (src/dext.cc:165-222). */                                    \
    size_t emitted = 0;                                                        \
    for (int32_t i = 0; i < ny; i++) {                                         \
      const int32_t i0 = i - KNL2 - 1;                                         \
      const int32_t i1 = i + KNL2 < ny ? i + KNL2 : ny - 1;                    \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        const T p = image[k] >= masked ? (T)0 : image[k];                      \
        if (!(p > 0 && mask1[k]))                                              \
          continue;                                                            \
        const int32_t j0 = j - KNL2 - 1;                                       \
        const int32_t j1 = j + KNL2 < nx ? j + KNL2 : nx - 1;                  \
        const W c = SAT_WINDOW(W, cnt, nx, i0, i1, j0, j1);                    \
        const W s = SAT_WINDOW(W, sum, nx, i0, i1, j0, j1);                    \
        const R mean = c >= 2 ? (R)s / (R)c : (R)0;                            \
        if (!((R)p >= mean + (R)3 * SQRT(mean)))                               \
          continue;                                                            \
        if (emitted < capacity) {                                              \
          oracle_signal_pixel *o = &out[emitted];                              \
          o->index = (uint32_t)k;                                              \
          o->value = (uint32_t)p;                                              \
          o->background = (float)mean;                                         \
          o->population = (uint16_t)c;                                         \
          o->reserved = 0;                                                     \
        }                                                                      \
        emitted++;                                                             \
      }                                                                        \
    }                                                                          \
    *found = emitted;                                                          \
    free(cnt); free(sum); free(sq); free(mask0); free(mask1);                  \
    return 0;                                                                  \
  }

/* R = float takes sqrtf and R = double takes sqrt, as std::sqrt's overloads
 * do in the reference; a float expression is never widened to double. */
DEFINE_DEXT(u16, uint16_t, uint32_t, float, 0xFFFFu, sqrtf)
DEFINE_DEXT(u32, uint32_t, uint64_t, double, 0xFFFFFFFFu, sqrt)

/* ------------------------------------------------------------------------ */
/* Threshold: direct clipped windows -- an independent second summation.    */
/* ------------------------------------------------------------------------ */

#define DEFINE_DEXT_DIRECT(SUFFIX, T, W, R, TMAX, SQRT)                        \
  int oracle_dext_direct_##SUFFIX(const T *image, size_t height,              \
                                  size_t width, oracle_signal_pixel *out,     \
                                  size_t capacity, size_t *found) {           \
    *found = 0;                                                                \
    if (image == NULL || height == 0 || width == 0)                            \
      return -1;                                                               \
    if (height > (size_t)INT32_MAX / width)                                    \
      return -1;                                                               \
    const int32_t ny = (int32_t)height, nx = (int32_t)width;                   \
    const size_t n = height * width;                                           \
    const T masked = (T)(TMAX - 1);                                            \
    uint8_t *m0 = (uint8_t *)malloc(n);                                        \
    uint8_t *m1 = (uint8_t *)malloc(n);                                        \
    if (!m0 || !m1) {                                                          \
      free(m0); free(m1);                                                      \
      return -4;                                                               \
    }                                                                          \
    for (int32_t i = 0; i < ny; i++)                                           \
      for (int32_t j = 0; j < nx; j++) {                                       \
        W c = 0, s = 0, q = 0;                                                 \
        for (int32_t a = i - KNL0; a <= i + KNL0; a++)                         \
          for (int32_t b = j - KNL0; b <= j + KNL0; b++) {                     \
            if (a < 0 || a >= ny || b < 0 || b >= nx)                          \
              continue;                                                        \
            const T v = image[(size_t)a * nx + b];                             \
            if (v >= masked)                                                   \
              continue;                                                        \
            c += 1;                                                            \
            s += (W)v;                                                         \
            q += (W)((W)v * (W)v);                                             \
          }                                                                    \
        int hit = 0;                                                           \
        if (c >= 2) {                                                          \
          const R rm = (R)c, rs = (R)s, rq = (R)q;                             \
          hit = (rm * rq - rs * rs - rs * (rm - (R)1)) >                       \
                (rs * (R)6 * SQRT((R)2 * (rm - (R)1)));                        \
        }                                                                      \
        const size_t k = (size_t)i * nx + j;                                   \
        m0[k] = (uint8_t)(hit || image[k] >= masked);                          \
      }                                                                        \
    for (int32_t i = 0; i < ny; i++)                                           \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        int all = 1;                                                           \
        for (int32_t a = i - ERODE; a <= i + ERODE; a++)                       \
          for (int32_t b = j - ERODE; b <= j + ERODE; b++)                     \
            if (a >= 0 && a < ny && b >= 0 && b < nx &&                        \
                !m0[(size_t)a * nx + b])                                       \
              all = 0;                                                         \
        m1[k] = (uint8_t)(image[k] >= masked || (m0[k] && all));               \
      }                                                                        \
    size_t emitted = 0;                                                        \
    for (int32_t i = 0; i < ny; i++)                                           \
      for (int32_t j = 0; j < nx; j++) {                                       \
        const size_t k = (size_t)i * nx + j;                                   \
        const T p = image[k] >= masked ? (T)0 : image[k];                      \
        if (!(p > 0 && m1[k]))                                                 \
          continue;                                                            \
        W c = 0, s = 0;                                                        \
        for (int32_t a = i - KNL2; a <= i + KNL2; a++)                         \
          for (int32_t b = j - KNL2; b <= j + KNL2; b++) {                     \
            if (a < 0 || a >= ny || b < 0 || b >= nx)                          \
              continue;                                                        \
            const size_t kk = (size_t)a * nx + b;                              \
            if (image[kk] >= masked || m1[kk])                                 \
              continue;                                                        \
            c += 1;                                                            \
            s += (W)image[kk];                                                 \
          }                                                                    \
        const R mean = c >= 2 ? (R)s / (R)c : (R)0;                            \
        if (!((R)p >= mean + (R)3 * SQRT(mean)))                               \
          continue;                                                            \
        if (emitted < capacity) {                                              \
          out[emitted].index = (uint32_t)k;                                    \
          out[emitted].value = (uint32_t)p;                                    \
          out[emitted].background = (float)mean;                               \
          out[emitted].population = (uint16_t)c;                               \
          out[emitted].reserved = 0;                                           \
        }                                                                      \
        emitted++;                                                             \
      }                                                                        \
    *found = emitted;                                                          \
    free(m0); free(m1);                                                        \
    return 0;                                                                  \
  }

DEFINE_DEXT_DIRECT(u16, uint16_t, uint32_t, float, 0xFFFFu, sqrtf)
DEFINE_DEXT_DIRECT(u32, uint32_t, uint64_t, double, 0xFFFFFFFFu, sqrt)

/* Many frames on host threads, for the CPU timing leg. */
typedef struct {
  const uint16_t *frames;
  size_t count, height, width;
  size_t next;
  pthread_mutex_t lock;
  long long total;
  int status;
} many_job;

static void *many_worker(void *arg) {
  many_job *job = (many_job *)arg;
  const size_t n = job->height * job->width;
  const size_t capacity = n / 8 + 1;
  oracle_signal_pixel *out =
      (oracle_signal_pixel *)malloc(capacity * sizeof(oracle_signal_pixel));
  for (;;) {
    pthread_mutex_lock(&job->lock);
    const size_t f = job->next++;
    pthread_mutex_unlock(&job->lock);
    if (f >= job->count)
      break;
    size_t found = 0;
    const int status = oracle_dext_u16(job->frames + f * n, job->height,
                                       job->width, out, capacity, &found);
    pthread_mutex_lock(&job->lock);
    if (status != 0)
      job->status = status;
    job->total += (long long)found;
    pthread_mutex_unlock(&job->lock);
  }
  free(out);
  return NULL;
}

long long oracle_dext_many_u16(const uint16_t *frames, size_t count,
                               size_t height, size_t width, int threads) {
  many_job job = {frames, count, height, width, 0,
                  PTHREAD_MUTEX_INITIALIZER, 0, 0};
  if (threads < 1)
    threads = 1;
  pthread_t *ids = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int t = 0; t < threads; t++)
    pthread_create(&ids[t], NULL, many_worker, &job);
  for (int t = 0; t < threads; t++)
    pthread_join(ids[t], NULL);
  free(ids);
  return job.status != 0 ? job.status : job.total;
}

/* ------------------------------------------------------------------------ */
/* Ordering: src/signal_order.cc:19-72. Ascending by index; duplicates keep   */
/* their input order (the contract the test pins against stable_sort,       */
/* tests/test_signal_order.cc:305-315). Restated as a stable counting sort   */
/* on the bucket (index >> shift) followed by an insertion sort per bucket.  */
/* ------------------------------------------------------------------------ */

static void insertion_sort(oracle_signal_pixel *p, size_t n) {
  for (size_t i = 1; i < n; i++) {
    oracle_signal_pixel x = p[i];
    size_t j = i;
    while (j > 0 && p[j - 1].index > x.index) {
      p[j] = p[j - 1];
      j--;
    }
    p[j] = x;
  }
}

void oracle_by_index(oracle_signal_pixel *pixels, size_t n) {
  if (n < 2)
    return;
  uint32_t largest = 0;
  for (size_t i = 0; i < n; i++)
    largest = pixels[i].index > largest ? pixels[i].index : largest;
  unsigned shift = 0;
  while ((largest >> shift) > n / 4 + 1)
    shift++;
  const size_t buckets = ((size_t)largest >> shift) + 1;
  size_t *start = (size_t *)calloc(buckets + 1, sizeof(size_t));
  oracle_signal_pixel *staged =
      (oracle_signal_pixel *)malloc(n * sizeof(oracle_signal_pixel));
  for (size_t i = 0; i < n; i++)
    start[(pixels[i].index >> shift) + 1]++;
  for (size_t b = 1; b <= buckets; b++)
    start[b] += start[b - 1];
  size_t *at = (size_t *)malloc(buckets * sizeof(size_t));
  memcpy(at, start, buckets * sizeof(size_t));
  for (size_t i = 0; i < n; i++)
    staged[at[pixels[i].index >> shift]++] = pixels[i];
  for (size_t b = 0; b < buckets; b++)
    insertion_sort(staged + start[b], start[b + 1] - start[b]);
  memcpy(pixels, staged, n * sizeof(oracle_signal_pixel));
  free(start);
  free(at);
  free(staged);
}

/* ------------------------------------------------------------------------ */
/* Grouping and centroids: src/dials_spots.cc, restated as a batch.          */
/*                                                                          */
/* Edges (src/dials_spots.cc:102-137): the pixel to the left in the same row  */
/* and frame (index % width != 0); the pixel above in the same frame; the    */
/* same index on the previous frame when the two frame numbers are           */
/* consecutive and not two_d. A gap in the numbering therefore connects      */
/* nothing across it, which is what flush(frame) at the gap achieves         */
/* (src/dials_spots.cc:68-69).                                              */
/*                                                                          */
/* Emission order: the streaming Labeller emits at each add() the components */
/* with no pixel on the frame just added, in order of first appearance in    */
/* its (frame, index)-ordered list (src/dials_spots.cc:153-194); every open  */
/* component then has its last pixel on that frame. So rows come out sorted  */
/* by (last frame, first (frame, index) member), which is what this sorts    */
/* by.                                                                      */
/* ------------------------------------------------------------------------ */

struct oracle_labeller {
  size_t height, width;
  size_t min_size, max_size;
  double max_separation;
  int two_d;
  int started;
  int64_t last_frame;
  /* every pixel handed in, in (frame, index) order */
  oracle_pixel *live;
  uint32_t *frame_of; /* position of the pixel's frame in add() order */
  size_t n, cap;
  size_t frames;
  size_t *frame_start; /* first pixel of each added frame; frames+1 entries */
  int64_t *frame_number;
  size_t frame_cap;
  /* results */
  oracle_spot *spots;
  size_t n_spots, spots_cap;
  oracle_pixel *kept;
  size_t n_kept, kept_cap;
  oracle_counts counts;
};

oracle_labeller *oracle_labeller_new(size_t height, size_t width,
                                     size_t min_spot_size,
                                     size_t max_spot_size,
                                     double max_separation, int two_d) {
  if (height == 0 || width == 0 || max_spot_size < min_spot_size ||
      min_spot_size == 0)
    return NULL; /* the three constructor throws, src/dials_spots.cc:22-30 */
  oracle_labeller *l = (oracle_labeller *)calloc(1, sizeof(oracle_labeller));
  l->height = height;
  l->width = width;
  l->min_size = min_spot_size;
  l->max_size = max_spot_size;
  l->max_separation = max_separation;
  l->two_d = two_d;
  return l;
}

void oracle_labeller_free(oracle_labeller *l) {
  if (!l)
    return;
  free(l->live);
  free(l->frame_of);
  free(l->frame_start);
  free(l->frame_number);
  free(l->spots);
  free(l->kept);
  free(l);
}

int oracle_labeller_add(oracle_labeller *l, int64_t frame,
                        const oracle_signal_pixel *signal, size_t n) {
  if (l->started && frame <= l->last_frame)
    return -1;
  const size_t pixels = l->height * l->width;
  for (size_t i = 0; i < n; i++) {
    if (signal[i].index >= pixels)
      return -3;
    if (i > 0 && signal[i].index <= signal[i - 1].index)
      return -2;
  }
  if (l->frames + 2 > l->frame_cap) {
    l->frame_cap = l->frame_cap ? 2 * l->frame_cap : 64;
    l->frame_start =
        (size_t *)realloc(l->frame_start, l->frame_cap * sizeof(size_t));
    l->frame_number =
        (int64_t *)realloc(l->frame_number, l->frame_cap * sizeof(int64_t));
  }
  if (l->n + n > l->cap) {
    while (l->n + n > l->cap)
      l->cap = l->cap ? 2 * l->cap : 1024;
    l->live = (oracle_pixel *)realloc(l->live, l->cap * sizeof(oracle_pixel));
    l->frame_of = (uint32_t *)realloc(l->frame_of, l->cap * sizeof(uint32_t));
  }
  l->frame_start[l->frames] = l->n;
  l->frame_number[l->frames] = frame;
  for (size_t i = 0; i < n; i++) {
    l->live[l->n + i].frame = (int32_t)frame;
    l->live[l->n + i].index = signal[i].index;
    l->live[l->n + i].value = signal[i].value;
    l->frame_of[l->n + i] = (uint32_t)l->frames;
  }
  l->n += n;
  l->frames++;
  l->frame_start[l->frames] = l->n;
  l->counts.signal_pixels += n;
  l->started = 1;
  l->last_frame = frame;
  return 0;
}

static uint32_t uf_find(uint32_t *parent, uint32_t i) {
  while (parent[i] != i) {
    parent[i] = parent[parent[i]];
    i = parent[i];
  }
  return i;
}

static void uf_union(uint32_t *parent, uint32_t a, uint32_t b) {
  a = uf_find(parent, a);
  b = uf_find(parent, b);
  if (a < b)
    parent[b] = a;
  else if (b < a)
    parent[a] = b;
}

/* Binary search for `index` among pixels [from, to) of one frame. */
static long find_index(const oracle_pixel *p, size_t from, size_t to,
                       uint32_t index) {
  size_t lo = from, hi = to;
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (p[mid].index < index)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < to && p[lo].index == index) ? (long)lo : -1;
}

static void push_spot(oracle_labeller *l, const oracle_spot *s) {
  if (l->n_spots == l->spots_cap) {
    l->spots_cap = l->spots_cap ? 2 * l->spots_cap : 64;
    l->spots =
        (oracle_spot *)realloc(l->spots, l->spots_cap * sizeof(oracle_spot));
  }
  l->spots[l->n_spots++] = *s;
}

static void push_kept(oracle_labeller *l, const oracle_pixel *p) {
  if (l->n_kept == l->kept_cap) {
    l->kept_cap = l->kept_cap ? 2 * l->kept_cap : 256;
    l->kept =
        (oracle_pixel *)realloc(l->kept, l->kept_cap * sizeof(oracle_pixel));
  }
  l->kept[l->n_kept++] = *p;
}

/* The per-component pass, src/dials_spots.cc:221-360: filters, bounding box,
 * sums in member order, centroid, variance, single-frame z, peak separation.
 */
static void emit_component(oracle_labeller *l, const uint32_t *members,
                           size_t n) {
  const double twelfth = 1.0 / 12.0;
  l->counts.groups++;
  if (n < l->min_size) {
    l->counts.too_small++;
    return;
  }
  if (n > l->max_size) {
    l->counts.too_large++;
    return;
  }
  oracle_spot spot;
  memset(&spot, 0, sizeof(spot));
  spot.n_signal = (uint32_t)n;
  double total = 0.0, total_sq = 0.0, moment[3] = {0.0, 0.0, 0.0};
  double peak[3] = {0.0, 0.0, 0.0};
  uint32_t peak_value = 0;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (size_t k = 0; k < n; k++) {
    const oracle_pixel *p = &l->live[members[k]];
    const int32_t c[3] = {(int32_t)(p->index % l->width),
                          (int32_t)(p->index / l->width), p->frame};
    for (int a = 0; a < 3; a++) {
      if (k == 0 || c[a] < lo[a])
        lo[a] = c[a];
      if (k == 0 || c[a] + 1 > hi[a])
        hi[a] = c[a] + 1;
    }
    const double v = (double)p->value;
    const double pos[3] = {(double)c[0] + 0.5, (double)c[1] + 0.5,
                           (double)c[2] + 0.5};
    total += v;
    total_sq += v * v;
    for (int a = 0; a < 3; a++)
      moment[a] += v * pos[a];
    if (p->value > peak_value) {
      peak_value = p->value;
      for (int a = 0; a < 3; a++)
        peak[a] = pos[a];
    }
  }
  for (int a = 0; a < 3; a++) {
    spot.bbox[2 * a] = lo[a];
    spot.bbox[2 * a + 1] = hi[a];
  }
  spot.intensity = total;
  spot.intensity_variance = total;
  if (total > 0.0) {
    for (int a = 0; a < 3; a++)
      spot.position[a] = moment[a] / total;
    double spread[3] = {0.0, 0.0, 0.0};
    for (size_t k = 0; k < n; k++) {
      const oracle_pixel *p = &l->live[members[k]];
      const double v = (double)p->value;
      const double pos[3] = {(double)(p->index % l->width) + 0.5,
                             (double)(p->index / l->width) + 0.5,
                             (double)p->frame + 0.5};
      for (int a = 0; a < 3; a++) {
        const double d = pos[a] - spot.position[a];
        spread[a] += v * d * d;
      }
    }
    if (total * total > total_sq) {
      for (int a = 0; a < 3; a++) {
        const double unbiased = spread[a] * total / (total * total - total_sq);
        spot.variance[a] = unbiased / total + twelfth;
      }
    } else {
      for (int a = 0; a < 3; a++)
        spot.variance[a] = twelfth;
    }
    if (spot.bbox[5] == spot.bbox[4] + 1)
      spot.position[2] = (double)spot.bbox[4] + 0.5;
  } else {
    for (int a = 0; a < 3; a++)
      spot.position[a] = (spot.bbox[2 * a + 1] + spot.bbox[2 * a]) / 2.0;
  }
  if (l->max_separation > 0.0) {
    const double dx = peak[0] - spot.position[0];
    const double dy = peak[1] - spot.position[1];
    const double dz = peak[2] - spot.position[2];
    if (sqrt(dx * dx + dy * dy + dz * dz) > l->max_separation) {
      l->counts.separated++;
      return;
    }
  }
  spot.first = l->n_kept;
  for (size_t k = 0; k < n; k++)
    push_kept(l, &l->live[members[k]]);
  push_spot(l, &spot);
  l->counts.accepted++;
}

int oracle_labeller_finish(oracle_labeller *l) {
  const size_t n = l->n;
  if (n == 0)
    return 0;
  const uint32_t w = (uint32_t)l->width;
  uint32_t *parent = (uint32_t *)malloc(n * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++)
    parent[i] = (uint32_t)i;
  for (size_t f = 0; f < l->frames; f++) {
    const size_t from = l->frame_start[f], to = l->frame_start[f + 1];
    const int linked = !l->two_d && f > 0 &&
                       l->frame_number[f - 1] + 1 == l->frame_number[f];
    for (size_t i = from; i < to; i++) {
      const uint32_t idx = l->live[i].index;
      if (i > from && idx % w != 0 && l->live[i - 1].index == idx - 1)
        uf_union(parent, (uint32_t)i, (uint32_t)(i - 1));
      if (idx >= w) {
        const long up = find_index(l->live, from, i, idx - w);
        if (up >= 0)
          uf_union(parent, (uint32_t)i, (uint32_t)up);
      }
      if (linked) {
        const long back =
            find_index(l->live, l->frame_start[f - 1], from, idx);
        if (back >= 0)
          uf_union(parent, (uint32_t)i, (uint32_t)back);
      }
    }
  }
  /* Roots are the smallest member (unions hook toward the lower position). */
  uint32_t *count = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
  uint32_t *last = (uint32_t *)malloc(n * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++) {
    const uint32_t r = uf_find(parent, (uint32_t)i);
    parent[i] = r;
    count[r + 1]++;
    last[r] = l->frame_of[i]; /* ascending i: ends at the last frame */
  }
  for (size_t i = 0; i < n; i++)
    count[i + 1] += count[i];
  uint32_t *member = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *cursor = (uint32_t *)malloc(n * sizeof(uint32_t));
  memcpy(cursor, count, n * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++)
    member[cursor[parent[i]]++] = (uint32_t)i;
  /* Components ordered by (last frame, root): a stable counting sort of the
   * roots (ascending already) by last frame. */
  const size_t frames = l->frames;
  uint32_t *by_frame = (uint32_t *)calloc(frames + 1, sizeof(uint32_t));
  size_t roots = 0;
  for (size_t i = 0; i < n; i++)
    if (parent[i] == i) {
      by_frame[last[i] + 1]++;
      roots++;
    }
  for (size_t f = 0; f < frames; f++)
    by_frame[f + 1] += by_frame[f];
  uint32_t *order = (uint32_t *)malloc(roots * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++)
    if (parent[i] == i)
      order[by_frame[last[i]]++] = (uint32_t)i;
  for (size_t c = 0; c < roots; c++) {
    const uint32_t r = order[c];
    emit_component(l, member + count[r], count[r + 1] - count[r]);
  }
  free(parent);
  free(count);
  free(last);
  free(member);
  free(cursor);
  free(by_frame);
  free(order);
  /* Everything is closed; further add() calls start afresh. */
  l->n = 0;
  l->frames = 0;
  return 0;
}

size_t oracle_labeller_spot_count(const oracle_labeller *l) {
  return l->n_spots;
}

size_t oracle_labeller_pixel_count(const oracle_labeller *l) {
  return l->n_kept;
}

void oracle_labeller_results(const oracle_labeller *l, oracle_spot *spots,
                             oracle_pixel *pixels, oracle_counts *counts) {
  if (spots && l->n_spots)
    memcpy(spots, l->spots, l->n_spots * sizeof(oracle_spot));
  if (pixels && l->n_kept)
    memcpy(pixels, l->kept, l->n_kept * sizeof(oracle_pixel));
  if (counts)
    *counts = l->counts;
}

/* ------------------------------------------------------------------------ */
/* Test frames: the generator the reference's device tests and benchmark     */
/* share, tests/synthetic_frame.hh:23-157, restated so that the parity tests */
/* run on the reference's own test frames. An LCG (multiplier                */
/* 6364136223846793005, increment 1442695040888963407, high 31 bits), Knuth  */
/* Poisson, a 5x5 weighted stamp clipped at the border, and optionally a     */
/* masked band at 0xfffe plus twenty 0xffff pixels.                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint64_t state;
} lcg;

static uint32_t lcg_next(lcg *r) {
  r->state = r->state * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(r->state >> 33);
}

static double lcg_uniform(lcg *r) { return (double)lcg_next(r) / 4294967296.0; }

static uint16_t lcg_counts(lcg *r, double mean) {
  const double limit = exp(-mean);
  double product = lcg_uniform(r);
  uint16_t k = 0;
  while (product > limit && k < 1000) {
    product *= lcg_uniform(r);
    k++;
  }
  return k;
}

static void stamp(uint16_t *px, size_t height, size_t width, size_t i,
                  size_t j, uint16_t peak) {
  static const int weight[5][5] = {{0, 1, 2, 1, 0},
                                   {1, 4, 8, 4, 1},
                                   {2, 8, 16, 8, 2},
                                   {1, 4, 8, 4, 1},
                                   {0, 1, 2, 1, 0}};
  for (int di = -2; di <= 2; di++)
    for (int dj = -2; dj <= 2; dj++) {
      const long a = (long)i + di, b = (long)j + dj;
      if (a < 0 || b < 0 || a >= (long)height || b >= (long)width)
        continue;
      const size_t k = (size_t)a * width + (size_t)b;
      const uint32_t v =
          px[k] + (uint32_t)peak * (uint32_t)weight[di + 2][dj + 2] / 16u;
      px[k] = (uint16_t)(v > 0xfffdu ? 0xfffdu : v);
    }
}

void oracle_synthetic_frame(uint16_t *px, size_t height, size_t width,
                            uint64_t seed, double background, int with_mask) {
  lcg r = {seed * 6364136223846793005ull + 1};
  const size_t area = height * width;
  for (size_t k = 0; k < area; k++)
    px[k] = lcg_counts(&r, background);
  if (height >= 48 && width >= 48) {
    stamp(px, height, width, 0, 0, 4000);
    stamp(px, height, width, 0, width / 2, 3000);
    stamp(px, height, width, height - 1, width - 1, 4000);
    stamp(px, height, width, height / 2, 0, 3500);
    stamp(px, height, width, height / 2, width - 1, 3500);
    stamp(px, height, width, 1, 1, 2500);
    stamp(px, height, width, height - 2, 2, 2500);
    const size_t scattered = area / 3000;
    for (size_t n = 0; n < scattered; n++) {
      const size_t i = 6 + lcg_next(&r) % (height - 13);
      const size_t j = 6 + lcg_next(&r) % (width - 13);
      stamp(px, height, width, i, j, (uint16_t)(800 + lcg_next(&r) % 6000));
    }
    stamp(px, height, width, height / 3, width / 3, 5000);
    stamp(px, height, width, height / 3, width / 3 + 2, 5000);
  } else {
    stamp(px, height, width, 0, 0, 4000);
    stamp(px, height, width, height - 1, width - 1, 4000);
  }
  if (with_mask) {
    for (size_t i = height / 4; i < height / 4 + 5 && i < height; i++)
      for (size_t j = 0; j < width; j++)
        px[i * width + j] = 0xfffe;
    for (int n = 0; n < 20; n++) {
      const size_t k = lcg_next(&r) % area;
      px[k] = 0xffff;
    }
  }
}
