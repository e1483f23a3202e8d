/*
 * Test infrastructure: writes NXmx-shaped HDF5 files the way a detector's
 * writer does -- data files whose image dataset holds one frame per chunk,
 * written with H5Dwrite_chunk (bytes as given, the declared filter never
 * run), and a master file whose /entry/data/data is a virtual dataset over
 * them -- so that the chunk reader can be checked against known bytes.
 * Loaded by tests/test_nxmx.py through ctypes.  0 on success, -1 on failure.
 */
#include <hdf5.h>
#include <stdint.h>
#include <string.h>

static hid_t type_of(unsigned bits) {
  return bits == 16 ? H5T_STD_U16LE : bits == 8 ? H5T_STD_U8LE : H5T_STD_U32LE;
}

static hid_t open_or_create(const char *path, int append) {
  if (append)
    return H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

static hid_t intermediate_groups(void) {
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  H5Pset_create_intermediate_group(lcpl, 1);
  return lcpl;
}

/* `frames` x h x w of `bits`-bit pixels in `dataset`, chunked
 * (chunk_frames, h, w), declaring `filter` (0 for none; declared optional,
 * since the plugin need not be installed); chunk i written from chunks[i]
 * (sizes[i] bytes, filter mask masks[i]) when sizes[i] > 0, else left
 * unallocated.  With append the file must exist and is added to. */
int nxw_write_data(const char *path, const char *dataset, int append,
                   uint64_t frames, uint64_t h, uint64_t w, unsigned bits,
                   unsigned filter, uint64_t chunk_frames,
                   const unsigned char *const *chunks, const uint64_t *sizes,
                   const unsigned *masks) {
  int rc = -1;
  hid_t file = open_or_create(path, append);
  if (file < 0)
    return -1;
  hsize_t dims[3] = {frames, h, w};
  hsize_t chunk[3] = {chunk_frames, h, w};
  hid_t space = H5Screate_simple(3, dims, NULL);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  hid_t lcpl = intermediate_groups();
  H5Pset_chunk(dcpl, 3, chunk);
  if (filter == 32008) {
    unsigned cd[5] = {0, 0, bits / 8, 0, 2}; /* bitshuffle: block 0 (default), lz4 */
    H5Pset_filter(dcpl, filter, H5Z_FLAG_OPTIONAL, 5, cd);
  } else if (filter != 0) {
    H5Pset_filter(dcpl, filter, H5Z_FLAG_OPTIONAL, 0, NULL);
  }
  hid_t dset = H5Dcreate2(file, dataset, type_of(bits), space, lcpl, dcpl, H5P_DEFAULT);
  if (dset >= 0) {
    rc = 0;
    for (uint64_t i = 0; chunks && i < frames / chunk_frames; i++) {
      if (sizes[i] == 0)
        continue;
      hsize_t off[3] = {i * chunk_frames, 0, 0};
      if (H5Dwrite_chunk(dset, H5P_DEFAULT, masks ? masks[i] : 0, off,
                         (size_t)sizes[i], chunks[i]) < 0)
        rc = -1;
    }
    H5Dclose(dset);
  }
  H5Pclose(lcpl);
  H5Pclose(dcpl);
  H5Sclose(space);
  H5Fclose(file);
  return rc;
}

/* A contiguous (not chunked) image dataset. */
int nxw_write_contiguous(const char *path, const char *dataset, uint64_t frames,
                         uint64_t h, uint64_t w) {
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    return -1;
  hsize_t dims[3] = {frames, h, w};
  hid_t space = H5Screate_simple(3, dims, NULL);
  hid_t lcpl = intermediate_groups();
  hid_t dset = H5Dcreate2(file, dataset, H5T_STD_U32LE, space, lcpl, H5P_DEFAULT,
                          H5P_DEFAULT);
  int rc = dset >= 0 ? 0 : -1;
  if (dset >= 0)
    H5Dclose(dset);
  H5Pclose(lcpl);
  H5Sclose(space);
  H5Fclose(file);
  return rc;
}

/* /entry/data/data in `path` as a virtual dataset of `images` x h x w over n
 * mappings: series frames firsts[k] .. firsts[k] + counts[k] - 1 are frames
 * offsets[k] .. of datasets[k] in files[k] ("." for this file itself); a
 * count of 0 maps an unlimited run.  With append the file must exist. */
int nxw_write_master(const char *path, int append, uint64_t images, uint64_t h,
                     uint64_t w, unsigned bits, size_t n,
                     const char *const *files, const char *const *datasets,
                     const uint64_t *firsts, const uint64_t *counts,
                     const uint64_t *offsets) {
  hid_t file = open_or_create(path, append);
  if (file < 0)
    return -1;
  hsize_t dims[3] = {images, h, w};
  hsize_t maxdims[3] = {H5S_UNLIMITED, h, w};
  hid_t vspace = H5Screate_simple(3, dims, maxdims);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  int rc = 0;
  for (size_t k = 0; k < n; k++) {
    hsize_t start[3] = {firsts[k], 0, 0};
    hsize_t count[3] = {counts[k] ? counts[k] : H5S_UNLIMITED, h, w};
    hsize_t sdims[3] = {offsets[k] + (counts[k] ? counts[k] : 1), h, w};
    hsize_t smax[3] = {H5S_UNLIMITED, h, w};
    hid_t src = H5Screate_simple(3, sdims, smax);
    hsize_t sstart[3] = {offsets[k], 0, 0};
    hsize_t one[3] = {1, 1, 1};
    if (counts[k]) {
      H5Sselect_hyperslab(vspace, H5S_SELECT_SET, start, NULL, count, NULL);
      H5Sselect_hyperslab(src, H5S_SELECT_SET, sstart, NULL, count, NULL);
    } else {
      hsize_t block[3] = {1, h, w};
      hsize_t cnt[3] = {H5S_UNLIMITED, 1, 1};
      H5Sselect_hyperslab(vspace, H5S_SELECT_SET, start, one, cnt, block);
      H5Sselect_hyperslab(src, H5S_SELECT_SET, sstart, one, cnt, block);
    }
    if (H5Pset_virtual(dcpl, vspace, files[k], datasets[k], src) < 0)
      rc = -1;
    H5Sclose(src);
  }
  H5Sselect_all(vspace);
  hid_t lcpl = intermediate_groups();
  hid_t dset = H5Dcreate2(file, "/entry/data/data", type_of(bits), vspace, lcpl,
                          dcpl, H5P_DEFAULT);
  if (dset < 0)
    rc = -1;
  else
    H5Dclose(dset);
  H5Pclose(lcpl);
  H5Pclose(dcpl);
  H5Sclose(vspace);
  H5Fclose(file);
  return rc;
}
