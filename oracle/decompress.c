/*
 * Oracle: decompression of one image, restating src/spots/decompress.cc:1-136
 * (decompress::image, decompress::frame_bytes) together with the two codecs it
 * calls.  Those live outside the reference tree, in the un-vendored submodule
 * third_party/bitshuffle (github.com/kiyo-masui/bitshuffle; the pinned commit
 * is not recorded in this copy) and the lz4 it bundles
 * (third_party/bitshuffle/lz4/lz4.c, cmake/spots.cmake:213-217):
 *
 *   oracle_lz4_decode_block   one LZ4 block (the lz4 block format), under
 *                             LZ4_decompress_safe's contract: the bytes written,
 *                             or a negative value for a malformed block;
 *   oracle_bitshuffle /       bitshuffle's bit transpose of one block of n
 *   oracle_bitunshuffle       elements (n a multiple of 8) and its inverse:
 *                             bit-row 8j + k holds bit k of byte j of every
 *                             element, element i at bit i % 8 of byte i / 8;
 *   oracle_bslz4_decode       bshuf_decompress_lz4: blocks of `block` elements,
 *                             the last rounded down to a multiple of 8, each a
 *                             4-byte big-endian compressed length and an LZ4
 *                             block of exactly block * elem bytes, then the
 *                             n % 8 leftover elements stored raw;
 *   oracle_decompress_image   decompress::image, with its checks and messages.
 *
 * Test infrastructure only: the product never links or calls it.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "oracle.h"

/* The block format's end-of-block rules, held as LZ4_decompress_safe holds
 * them: the last sequence is literals only and ends the input; any other
 * sequence leaves at least 8 input bytes after its literals (an offset, and a
 * last sequence of a token and 5 literals); a match starts at least 12 bytes
 * and ends at least 5 bytes before the output capacity's end.  Offset 0, or
 * one reaching before the block's first byte, is malformed. */
enum { kMfLimit = 12, kLastLiterals = 5, kMinMatch = 4 };

long oracle_lz4_decode_block(const uint8_t *src, size_t n, uint8_t *dst,
                             size_t cap) {
  size_t ip = 0, op = 0;
  if (n == 0)
    return -1;
  for (;;) {
    if (ip >= n)
      return -1;
    const unsigned token = src[ip++];
    size_t lit = token >> 4;
    if (lit == 15) {
      unsigned b;
      do {
        if (ip >= n)
          return -1;
        b = src[ip++];
        lit += b;
      } while (b == 255);
    }
    if (lit > n - ip || lit > cap - op)
      return -1;
    memcpy(dst + op, src + ip, lit);
    ip += lit;
    op += lit;
    if (ip == n)
      return (long)op; /* the last sequence holds literals only */
    if (cap < kMfLimit || op > cap - kMfLimit || n - ip < 8)
      return -1; /* not the last sequence, yet too near either end */
    const size_t offset = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
    ip += 2;
    size_t len = token & 15;
    if (len == 15) {
      unsigned b;
      do {
        if (ip >= n)
          return -1;
        b = src[ip++];
        len += b;
      } while (b == 255);
    }
    len += kMinMatch;
    if (offset == 0 || offset > op)
      return -1;
    if (len > cap - kLastLiterals - op)
      return -1;
    /* byte by byte, so that an overlapping match repeats its period */
    for (size_t i = 0; i < len; i++)
      dst[op + i] = dst[op - offset + i];
    op += len;
  }
}

void oracle_bitshuffle(const uint8_t *in, uint8_t *out, size_t n, size_t elem) {
  const size_t row = n / 8; /* bytes per bit-row */
  memset(out, 0, n * elem);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < elem; j++)
      for (size_t k = 0; k < 8; k++)
        if ((in[i * elem + j] >> k) & 1u)
          out[(8 * j + k) * row + i / 8] |= (uint8_t)(1u << (i % 8));
}

void oracle_bitunshuffle(const uint8_t *in, uint8_t *out, size_t n,
                         size_t elem) {
  const size_t row = n / 8;
  memset(out, 0, n * elem);
  for (size_t j = 0; j < elem; j++)
    for (size_t k = 0; k < 8; k++)
      for (size_t i = 0; i < n; i++)
        if ((in[(8 * j + k) * row + i / 8] >> (i % 8)) & 1u)
          out[i * elem + j] |= (uint8_t)(1u << k);
}

static uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) |
         ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

/* bitshuffle's default block: 8192 bytes of elements, a multiple of 8, at
 * least 128 (bshuf_default_block_size). */
size_t oracle_bslz4_default_block(size_t elem) {
  size_t b = (8192 / elem) / 8 * 8;
  return b < 128 ? 128 : b;
}

/* One block of bshuf_decompress_lz4: bytes consumed (4 + the compressed
 * length), or a negative code: -91 when the block does not decompress to
 * exactly count * elem bytes, the LZ4 code - 1000 when it is malformed, and
 * -92 (ours: bitshuffle takes no input length) when the length runs past the
 * chunk. */
static long bslz4_block(const uint8_t *in, size_t avail, uint8_t *out,
                        size_t count, size_t elem, uint8_t *tmp) {
  if (avail < 4)
    return -92;
  const uint32_t c = be32(in);
  if (c > avail - 4)
    return -92;
  const long got = oracle_lz4_decode_block(in + 4, c, tmp, count * elem);
  if (got < 0)
    return got - 1000;
  if ((size_t)got != count * elem)
    return -91;
  oracle_bitunshuffle(tmp, out, count, elem);
  return 4 + (long)c;
}

long oracle_bslz4_decode(const uint8_t *in, size_t n, uint8_t *out,
                         size_t count, size_t elem, size_t block) {
  if (block == 0)
    block = oracle_bslz4_default_block(elem);
  if (block % 8 != 0)
    return -81;
  static uint8_t tmp[1u << 24];
  if (block * elem > sizeof tmp)
    return -1;
  size_t ip = 0, done = 0;
  long err = 0;
  const size_t full = count / block;
  for (size_t b = 0; b < full; b++, done += block) {
    const long r = bslz4_block(in + ip, n - ip, out + done * elem, block, elem, tmp);
    if (r < 0) {
      err = r;
      if (r == -92)
        return r; /* nothing past a length that runs off the chunk */
      /* bitshuffle carries on past a bad block, to the next length */
      const uint32_t c = be32(in + ip);
      ip += 4 + c;
      continue;
    }
    ip += (size_t)r;
  }
  const size_t last = (count % block) / 8 * 8;
  if (last) {
    const long r = bslz4_block(in + ip, n - ip, out + done * elem, last, elem, tmp);
    if (r < 0)
      return r;
    ip += (size_t)r;
    done += last;
  }
  if (err < 0)
    return err;
  const size_t left = (count % 8) * elem;
  if (left > n - ip)
    return -92;
  memcpy(out + done * elem, in + ip, left);
  return (long)(ip + left);
}

static int fail(char *err, size_t len, const char *msg) {
  if (err && len)
    snprintf(err, len, "%s", msg);
  return -1;
}

int oracle_decompress_image(const uint8_t *data, size_t n, const char *algorithm,
                            unsigned bit_depth, size_t height, size_t width,
                            uint8_t *out, size_t out_size, char *err,
                            size_t err_len) {
  char msg[256];
  const size_t expected = height * width * bit_depth / 8;
  const size_t element = bit_depth / 8;
  if (element == 0 || out_size < expected)
    return fail(err, err_len, "output buffer too small for the frame");
  if (algorithm == NULL || algorithm[0] == '\0') {
    if (n != expected) {
      snprintf(msg, sizeof msg, "uncompressed frame is %zu bytes, expected %zu",
               n, expected);
      return fail(err, err_len, msg);
    }
    memcpy(out, data, expected);
    return 0;
  }
  if (n < 12)
    return fail(err, err_len, "too short for a filter header");
  uint64_t size = 0;
  for (int i = 0; i < 8; i++)
    size = (size << 8) | data[i];
  const uint32_t block = be32(data + 8);
  if (size != expected) {
    snprintf(msg, sizeof msg,
             "filter header says %llu bytes uncompressed, expected %zu",
             (unsigned long long)size, expected);
    return fail(err, err_len, msg);
  }
  const uint8_t *payload = data + 12;
  const size_t plen = n - 12;
  if (strcmp(algorithm, "bslz4") == 0) {
    const long r = oracle_bslz4_decode(payload, plen, out, expected / element,
                                       element, block / element);
    if (r < 0) {
      snprintf(msg, sizeof msg, "bitshuffle decompression failed with code %ld", r);
      return fail(err, err_len, msg);
    }
    return 0;
  }
  if (strcmp(algorithm, "lz4") == 0) {
    /* decompress.cc:43-72: each block a 4-byte big-endian length and an LZ4
     * block of at most the block size, until the frame is full */
    const size_t bs = block > 0 ? block : expected;
    size_t read = 0, written = 0;
    while (written < expected) {
      if (read + 4 > plen)
        return fail(err, err_len, "lz4 data truncated");
      const uint32_t c = be32(payload + read);
      read += 4;
      if (c == 0 || read + c > plen)
        return fail(err, err_len, "lz4 block size out of range");
      const size_t remaining = expected - written;
      const size_t want = remaining < bs ? remaining : bs;
      const long got = oracle_lz4_decode_block(payload + read, c, out + written, want);
      if (got < 0) {
        snprintf(msg, sizeof msg, "lz4 decompression failed with code %ld", got);
        return fail(err, err_len, msg);
      }
      read += c;
      written += (size_t)got;
    }
    return 0;
  }
  snprintf(msg, sizeof msg, "unknown compression algorithm %s", algorithm);
  return fail(err, err_len, msg);
}
