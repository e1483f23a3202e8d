/*
 * The oracle's grouping: a CPU restatement of dials_spots::Labeller
 * (src/spots/dials_spots.cc:36-360, dials_spots.hh:108-158), for the checker
 * only.
 *
 * Restated behaviour:
 *   add      frames strictly ascending, or an error      [dials_spots.cc:57-63]
 *            a gap in frame numbers closes everything    [:68-69]
 *            pixels ascending and inside the frame       [:80-93]
 *            union left (not at column 0), up (index-W) and, in 3D, the same
 *            index on the previous frame, toward the lower position [:104-137]
 *            then flush(two_d ? frame + 1 : frame)      [:144]
 *   flush    labels in order of first appearance, counting sort by label,
 *            emit every component with no pixel on keep_frame, compact the
 *            rest in (frame, index) order                [:153-219]
 *   emit     size filters, bbox, sums, centroid at pixel centres, unbiased
 *            variance + 1/12 (1/12 alone when sum^2 <= sum of squares), a
 *            one-frame spot at z0 + 0.5, peak = first strictly greater value,
 *            separation filter                           [:221-360]
 *
 * One difference that no caller can see: add() validates the whole frame
 * before changing any state, where the reference throws part way through an
 * append.  Both refuse the frame; what the reference holds after such a throw
 * is not specified by it.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE 0xFFFFFFFFu

typedef struct {
  void *data;
  size_t size, cap, elem;
} vec;

static void vec_reserve(vec *v, size_t n) {
  if (n <= v->cap)
    return;
  size_t cap = v->cap ? v->cap : 64;
  while (cap < n)
    cap *= 2;
  v->data = realloc(v->data, cap * v->elem);
  v->cap = cap;
}

static void vec_resize(vec *v, size_t n) {
  vec_reserve(v, n);
  v->size = n;
}

#define AT(v, T, i) (((T *)(v).data)[i])

struct oracle_labeller {
  size_t height, width;
  size_t min_spot_size, max_spot_size;
  double max_separation;
  int two_d;

  vec live;   /* oracle_pixel */
  vec parent; /* uint32_t */
  size_t previous;
  int64_t last_frame;
  int started;

  vec spots; /* oracle_spot */
  vec kept;  /* oracle_pixel */
  oracle_counts counts;

  vec root, label, offset, cursor, member; /* uint32_t */
  vec open;                                /* char */
};

static void set_err(char *err, size_t errlen, const char *msg) {
  if (err && errlen)
    snprintf(err, errlen, "%s", msg);
}

oracle_labeller *oracle_labeller_new(size_t height, size_t width,
                                     size_t min_spot_size, size_t max_spot_size,
                                     double max_separation, int two_d,
                                     char *err, size_t errlen) {
  if (height == 0 || width == 0) {
    set_err(err, errlen, "dials_spots: a frame with no pixels in it");
    return NULL;
  }
  if (max_spot_size < min_spot_size) {
    set_err(err, errlen,
            "dials_spots: max_spot_size is below min_spot_size, so nothing "
            "can be accepted");
    return NULL;
  }
  if (min_spot_size == 0) {
    set_err(err, errlen, "dials_spots: min_spot_size must be at least 1");
    return NULL;
  }
  oracle_labeller *l = (oracle_labeller *)calloc(1, sizeof(*l));
  l->height = height;
  l->width = width;
  l->min_spot_size = min_spot_size;
  l->max_spot_size = max_spot_size;
  l->max_separation = max_separation;
  l->two_d = two_d;
  l->live.elem = sizeof(oracle_pixel);
  l->parent.elem = sizeof(uint32_t);
  l->spots.elem = sizeof(oracle_spot);
  l->kept.elem = sizeof(oracle_pixel);
  l->root.elem = l->label.elem = l->offset.elem = l->cursor.elem =
      l->member.elem = sizeof(uint32_t);
  l->open.elem = 1;
  return l;
}

void oracle_labeller_free(oracle_labeller *l) {
  if (!l)
    return;
  vec *all[] = {&l->live,  &l->parent, &l->spots,  &l->kept,   &l->root,
                &l->label, &l->offset, &l->cursor, &l->member, &l->open};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++)
    free(all[i]->data);
  free(l);
}

static uint32_t find_root(oracle_labeller *l, uint32_t i) {
  uint32_t *parent = (uint32_t *)l->parent.data;
  while (parent[i] != i) {
    parent[i] = parent[parent[i]];
    i = parent[i];
  }
  return i;
}

static void merge(oracle_labeller *l, uint32_t a, uint32_t b) {
  const uint32_t ra = find_root(l, a), rb = find_root(l, b);
  if (ra == rb)
    return;
  uint32_t *parent = (uint32_t *)l->parent.data;
  if (ra < rb)
    parent[rb] = ra;
  else
    parent[ra] = rb;
}

static void emit(oracle_labeller *l, const uint32_t *members, size_t n) {
  l->counts.groups++;
  if (n < l->min_spot_size) {
    l->counts.too_small++;
    return;
  }
  if (n > l->max_spot_size) {
    l->counts.too_large++;
    return;
  }
  const oracle_pixel *live = (const oracle_pixel *)l->live.data;
  oracle_spot spot;
  memset(&spot, 0, sizeof(spot));
  spot.n_signal = (uint32_t)n;

  double sum = 0.0, sum_sq = 0.0, weighted[3] = {0.0, 0.0, 0.0};
  int32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  uint32_t peak_value = 0;
  double peak[3] = {0.0, 0.0, 0.0};
  for (size_t k = 0; k < n; k++) {
    const oracle_pixel *p = &live[members[k]];
    const int32_t x = (int32_t)(p->index % l->width);
    const int32_t y = (int32_t)(p->index / l->width);
    const int32_t z = p->frame;
    if (k == 0) {
      x0 = x, x1 = x + 1, y0 = y, y1 = y + 1, z0 = z, z1 = z + 1;
    } else {
      x0 = x < x0 ? x : x0;
      x1 = x + 1 > x1 ? x + 1 : x1;
      y0 = y < y0 ? y : y0;
      y1 = y + 1 > y1 ? y + 1 : y1;
      z0 = z < z0 ? z : z0;
      z1 = z + 1 > z1 ? z + 1 : z1;
    }
    const double value = (double)p->value;
    const double pos[3] = {(double)x + 0.5, (double)y + 0.5, (double)z + 0.5};
    sum += value;
    sum_sq += value * value;
    for (int a = 0; a < 3; a++)
      weighted[a] += value * pos[a];
    if (p->value > peak_value) {
      peak_value = p->value;
      peak[0] = pos[0], peak[1] = pos[1], peak[2] = pos[2];
    }
  }
  spot.bbox[0] = x0, spot.bbox[1] = x1, spot.bbox[2] = y0;
  spot.bbox[3] = y1, spot.bbox[4] = z0, spot.bbox[5] = z1;
  spot.intensity = sum;
  spot.intensity_variance = sum;
  if (sum > 0.0) {
    for (int a = 0; a < 3; a++)
      spot.position[a] = weighted[a] / sum;
    double delta_sq[3] = {0.0, 0.0, 0.0};
    for (size_t k = 0; k < n; k++) {
      const oracle_pixel *p = &live[members[k]];
      const double value = (double)p->value;
      const double pos[3] = {(double)(p->index % l->width) + 0.5,
                             (double)(p->index / l->width) + 0.5,
                             (double)p->frame + 0.5};
      for (int a = 0; a < 3; a++) {
        const double d = pos[a] - spot.position[a];
        delta_sq[a] += value * d * d;
      }
    }
    if (sum * sum > sum_sq) {
      for (int a = 0; a < 3; a++) {
        const double v = delta_sq[a] * sum / (sum * sum - sum_sq);
        spot.variance[a] = v / sum + 1.0 / 12.0;
      }
    } else {
      for (int a = 0; a < 3; a++)
        spot.variance[a] = 1.0 / 12.0;
    }
    if (spot.bbox[5] == spot.bbox[4] + 1)
      spot.position[2] = (double)spot.bbox[4] + 0.5;
  } else {
    spot.position[0] = (spot.bbox[1] + spot.bbox[0]) / 2.0;
    spot.position[1] = (spot.bbox[3] + spot.bbox[2]) / 2.0;
    spot.position[2] = (spot.bbox[5] + spot.bbox[4]) / 2.0;
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
  spot.first = l->kept.size;
  vec_reserve(&l->kept, l->kept.size + n);
  for (size_t k = 0; k < n; k++)
    AT(l->kept, oracle_pixel, l->kept.size + k) = live[members[k]];
  l->kept.size += n;
  vec_reserve(&l->spots, l->spots.size + 1);
  AT(l->spots, oracle_spot, l->spots.size) = spot;
  l->spots.size++;
  l->counts.accepted++;
}

static void flush(oracle_labeller *l, int64_t keep_frame) {
  const size_t n = l->live.size;
  if (n == 0) {
    l->previous = 0;
    return;
  }
  vec_resize(&l->root, n);
  vec_resize(&l->label, n);
  uint32_t *root = (uint32_t *)l->root.data;
  uint32_t *label = (uint32_t *)l->label.data;
  for (size_t i = 0; i < n; i++)
    label[i] = NONE;
  uint32_t components = 0;
  for (size_t i = 0; i < n; i++) {
    const uint32_t r = find_root(l, (uint32_t)i);
    if (label[r] == NONE)
      label[r] = components++;
    root[i] = label[r];
  }
  vec_resize(&l->offset, (size_t)components + 1);
  uint32_t *offset = (uint32_t *)l->offset.data;
  memset(offset, 0, sizeof(uint32_t) * ((size_t)components + 1));
  for (size_t i = 0; i < n; i++)
    offset[root[i] + 1]++;
  for (size_t c = 0; c < components; c++)
    offset[c + 1] += offset[c];
  vec_resize(&l->cursor, components > 0 ? components : 1);
  uint32_t *cursor = (uint32_t *)l->cursor.data;
  memcpy(cursor, offset, sizeof(uint32_t) * components);
  vec_resize(&l->member, n);
  uint32_t *member = (uint32_t *)l->member.data;
  for (size_t i = 0; i < n; i++)
    member[cursor[root[i]]++] = (uint32_t)i;

  vec_resize(&l->open, components > 0 ? components : 1);
  char *open = (char *)l->open.data;
  memset(open, 0, components);
  const oracle_pixel *live = (const oracle_pixel *)l->live.data;
  for (size_t i = 0; i < n; i++)
    if (live[i].frame == keep_frame)
      open[root[i]] = 1;

  for (uint32_t c = 0; c < components; c++)
    if (!open[c])
      emit(l, &member[offset[c]], offset[c + 1] - offset[c]);

  /* emit() may have grown nothing of ours, but refresh pointers anyway */
  root = (uint32_t *)l->root.data;
  cursor = (uint32_t *)l->cursor.data;
  open = (char *)l->open.data;
  oracle_pixel *wlive = (oracle_pixel *)l->live.data;
  uint32_t *parent = (uint32_t *)l->parent.data;
  for (uint32_t c = 0; c < components; c++)
    cursor[c] = NONE;
  size_t write = 0;
  for (size_t i = 0; i < n; i++) {
    const uint32_t c = root[i];
    if (!open[c])
      continue;
    if (cursor[c] == NONE)
      cursor[c] = (uint32_t)write;
    wlive[write] = wlive[i];
    parent[write] = cursor[c];
    write++;
  }
  l->live.size = write;
  l->parent.size = write;
  l->previous = write;
  while (l->previous > 0 && wlive[l->previous - 1].frame == keep_frame)
    l->previous--;
}

int oracle_labeller_add(oracle_labeller *l, int64_t frame,
                        const oracle_signal_pixel *signal, size_t n, char *err,
                        size_t errlen) {
  char msg[256];
  if (l->started && frame <= l->last_frame) {
    snprintf(msg, sizeof msg,
             "dials_spots: frame %lld arrived after frame %lld; frames must be "
             "added in ascending order",
             (long long)frame, (long long)l->last_frame);
    set_err(err, errlen, msg);
    return -1;
  }
  const size_t pixels = l->height * l->width;
  for (size_t i = 0; i < n; i++) {
    if (signal[i].index >= pixels) {
      snprintf(msg, sizeof msg,
               "dials_spots: signal pixel %u is off the end of the frame",
               signal[i].index);
      set_err(err, errlen, msg);
      return -1;
    }
    if (i > 0 && signal[i].index <= signal[i - 1].index) {
      set_err(err, errlen,
              "dials_spots: signal pixels are not ascending by index");
      return -1;
    }
  }
  if (l->started && frame != l->last_frame + 1)
    flush(l, frame);

  const size_t base = l->live.size;
  if (base + n > (size_t)NONE) {
    set_err(err, errlen,
            "dials_spots: more than 4 billion pixels are held open at once");
    return -1;
  }
  vec_resize(&l->live, base + n);
  vec_resize(&l->parent, base + n);
  oracle_pixel *live = (oracle_pixel *)l->live.data;
  uint32_t *parent = (uint32_t *)l->parent.data;
  for (size_t i = 0; i < n; i++) {
    live[base + i].frame = (int32_t)frame;
    live[base + i].index = signal[i].index;
    live[base + i].value = signal[i].value;
    parent[base + i] = (uint32_t)(base + i);
  }
  l->counts.signal_pixels += n;
  const size_t end = base + n;
  const uint32_t columns = (uint32_t)l->width;

  for (size_t i = base + 1; i < end; i++)
    if (live[i].index % columns != 0 && live[i].index == live[i - 1].index + 1)
      merge(l, (uint32_t)i, (uint32_t)(i - 1));

  for (size_t i = base, up = base; i < end; i++) {
    if (live[i].index < columns)
      continue;
    const uint32_t target = live[i].index - columns;
    while (up < i && live[up].index < target)
      up++;
    if (up < i && live[up].index == target)
      merge(l, (uint32_t)i, (uint32_t)up);
  }

  if (!l->two_d && l->started && l->last_frame + 1 == frame &&
      l->previous < base) {
    for (size_t i = base, back = l->previous; i < end; i++) {
      const uint32_t target = live[i].index;
      while (back < base && live[back].index < target)
        back++;
      if (back < base && live[back].index == target)
        merge(l, (uint32_t)i, (uint32_t)back);
    }
  }
  l->started = 1;
  l->last_frame = frame;
  flush(l, l->two_d ? frame + 1 : frame);
  return 0;
}

void oracle_labeller_finish(oracle_labeller *l) {
  if (!l->started)
    return;
  flush(l, l->last_frame + 1);
}

size_t oracle_labeller_spot_count(const oracle_labeller *l) {
  return l->spots.size;
}
const oracle_spot *oracle_labeller_spots(const oracle_labeller *l) {
  return (const oracle_spot *)l->spots.data;
}
size_t oracle_labeller_pixel_count(const oracle_labeller *l) {
  return l->kept.size;
}
const oracle_pixel *oracle_labeller_pixels(const oracle_labeller *l) {
  return (const oracle_pixel *)l->kept.data;
}
oracle_counts oracle_labeller_counts(const oracle_labeller *l) {
  return l->counts;
}
size_t oracle_labeller_live(const oracle_labeller *l) { return l->live.size; }
