/*
 * The CPU baseline driver for the oracle's threshold: N worker threads, each
 * taking the next frame and thresholding it alone, which is how mxi_find -j N
 * spreads frames (src/spots/find_spots.cc:732-835).  Test infrastructure only.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <time.h>

#include "oracle.h"

typedef struct {
  const uint32_t *frames;
  size_t nframes, total, height, width;
  atomic_size_t next;
  atomic_uint_fast64_t signal;
} job;

static void *worker(void *arg) {
  job *j = (job *)arg;
  const size_t n = j->height * j->width;
  const size_t capacity = n / 4;
  oracle_signal_pixel *out =
      (oracle_signal_pixel *)malloc(sizeof(oracle_signal_pixel) * capacity);
  for (;;) {
    const size_t f = atomic_fetch_add(&j->next, 1);
    if (f >= j->total)
      break;
    size_t found = 0;
    oracle_dext_u32(j->frames + (f % j->nframes) * n, j->height, j->width, out,
                    capacity, &found);
    atomic_fetch_add(&j->signal, (uint_fast64_t)found);
  }
  free(out);
  return NULL;
}

uint64_t oracle_dext_u32_threads(const uint32_t *frames, size_t nframes,
                                 size_t total, size_t height, size_t width,
                                 int threads, double *seconds) {
  job j;
  j.frames = frames;
  j.nframes = nframes;
  j.total = total;
  j.height = height;
  j.width = width;
  atomic_init(&j.next, 0);
  atomic_init(&j.signal, 0);
  if (threads < 1)
    threads = 1;
  pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int i = 0; i < threads; i++)
    pthread_create(&t[i], NULL, worker, &j);
  for (int i = 0; i < threads; i++)
    pthread_join(t[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &b);
  free(t);
  if (seconds)
    *seconds = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
  return (uint64_t)atomic_load(&j.signal);
}
