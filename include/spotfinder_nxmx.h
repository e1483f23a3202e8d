/*
 * spotfinder_nxmx: an NXmx series' frames as raw HDF5 chunks, for the device
 * decoder.  The C ABI of libspotfinder_nxmx.so, the reader the reference's
 * series::nxmx provides (src/spots/series.hh:32-97, src/spots/nxmx.cc):
 * the master file's /entry/data/data, a virtual dataset over data files or a
 * plain chunked dataset, is resolved into blocks of frames, and each frame is
 * read with H5Dread_chunk straight out of the file that holds it, so that its
 * compressed bytes reach the decoder without HDF5's filter pipeline running
 * (and without the bitshuffle plugin installed).
 *
 * A separate library from libspotfinder_hip.so, so that the device backend
 * itself never needs HDF5.  Every HDF5 call is made under one process-wide
 * lock (nxmx.cc:41-59): a series may be shared by threads, and one series a
 * thread works as well.
 *
 * Status codes: 0 success; 1 for a frame the writer never stored (an
 * unallocated chunk, Reader::read's false) or a master file that is not there
 * yet (Series::try_open's false); -1 bad argument (a null pointer, a frame
 * outside every mapping); -2 a buffer too small (the size needed is
 * returned); -3 anything HDF5 or the file refuses, with the message in
 * sf_nxmx_last_error().
 */
#ifndef SPOTFINDER_NXMX_H
#define SPOTFINDER_NXMX_H

#include <stddef.h>
#include <stdint.h>

#include "spotfinder_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_nxmx sf_nxmx;

/* series::Info (series.hh:59-64), less the name. */
typedef struct sf_nxmx_info {
  uint64_t images; /* frames in /entry/data/data */
  uint64_t height; /* slow */
  uint64_t width;  /* fast */
} sf_nxmx_info;

/* series::SourceBlock (series.hh:101-104): frames first..last (inclusive) of
 * the series are frames offset.. of `dataset` in `filename` (resolved against
 * the master's directory; "." in the mapping is the master itself).  The
 * strings belong to the series. */
typedef struct sf_nxmx_block {
  const char *filename;
  const char *dataset;
  uint64_t first, last, offset;
} sf_nxmx_block;

/* series::Frame's description (series.hh:32-45): the frame's number in the
 * series, its shape and pixel width, the chunk's size in bytes and its
 * algorithm ("bslz4", "lz4", or "" for a chunk stored uncompressed; owned by
 * the series). */
typedef struct sf_nxmx_frame {
  uint64_t number;
  uint64_t height, width;
  uint32_t bit_depth;
  const char *algorithm;
  size_t size;
} sf_nxmx_frame;

/* Series::try_open + reader (nxmx.cc:330-386, 433-497): 0 and the series, 1
 * when `master` is not (yet) a regular file, -3 with a message otherwise (not
 * HDF5, JSON given in the images' place, not a stack of images, a mapping to
 * nothing). */
int sf_nxmx_open(const char *master, sf_nxmx **out, sf_nxmx_info *info);
void sf_nxmx_close(sf_nxmx *s);

/* The resolved blocks in ascending frame order (nxmx_blocks, nxmx.cc:519-530). */
size_t sf_nxmx_block_count(const sf_nxmx *s);
int sf_nxmx_get_block(const sf_nxmx *s, size_t i, sf_nxmx_block *b);

/* A frame's chunk, without reading it: 0 (and its description), 1 absent. */
int sf_nxmx_frame_info(sf_nxmx *s, uint64_t frame, sf_nxmx_frame *f);

/* Reader::read (nxmx.cc:194-236): the frame's chunk, bytes exactly as HDF5
 * stored them, into `buf` (`capacity` bytes).  0; 1 absent; -2 when the
 * chunk is larger than `capacity` (f->size says how large). */
int sf_nxmx_read(sf_nxmx *s, uint64_t frame, void *buf, size_t capacity,
                 sf_nxmx_frame *f);

/* Frames first .. first + n - 1 read back to back into `buf` (`capacity`
 * bytes), each present one described in chunks[k] for sf_decoder_load, in
 * frame order, with its device frame device_frames[i] (i the frame's place
 * in the range).  *loaded is the number of chunks described, *used the
 * bytes taken; present[i] (optional) 1 or 0.  -2 when `buf` is too small
 * (*used is what the range needs). */
int sf_nxmx_read_chunks(sf_nxmx *s, uint64_t first, size_t n, uint8_t *buf,
                        size_t capacity, void *const *device_frames,
                        sf_chunk *chunks, size_t *loaded, size_t *used,
                        int *present);

/* The calling thread's last failure. */
const char *sf_nxmx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
