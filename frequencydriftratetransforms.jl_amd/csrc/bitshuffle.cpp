// Decoder for bitshuffle-compressed HDF5 chunks (filter 32008), host side.
//
// The reference reads its spectrograms from Breakthrough Listen filterbank
// files, whose chunks are compressed with the bitshuffle HDF5 filter: its
// heavy test loads the H5Zbitshuffle plugin package for that
// (test/heavytests.jl:6).  This decoder lets the ingest read such chunks
// through HDF5's direct chunk reads when no plugin is installed.
//
// Format, as the bitshuffle library's HDF5 filter writes it (filter
// parameters cd[2] = element size, cd[3] = block size in elements, 0 for the
// default, cd[4] = 2 for LZ4):
//   * LZ4: a 12-byte header, the uncompressed size in bytes (uint64) and the
//     block size in bytes (uint32), both big-endian; then, per block of
//     block-size elements (the last one cut to a multiple of 8 elements), the
//     block's compressed size (uint32, big-endian) and the LZ4 block of its
//     bit-transposed bytes; the size mod 8 leftover elements follow raw.
//   * no compression: the blocks' bit-transposed bytes back to back, the
//     leftover elements raw (no header).
// Bit transposition of a block of n elements (n a multiple of 8) of E bytes:
// the output is 8E planes of n/8 bytes; plane 8b + k holds bit k of byte b of
// every element, element i at bit i mod 8 of byte i/8.
//
// The LZ4 block decoder follows the LZ4 block format description (token:
// literal length and match length - 4 in 4 bits each, 255-continued; literals;
// a 2-byte little-endian match offset; the last sequence holds only
// literals).  Every read and write is bounds-checked.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "zdt_internal.h"

namespace zdt {
namespace {

uint64_t be64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
  return v;
}
uint32_t be32(const uint8_t* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

// LZ4 block into out[0, cap); returns the bytes written, or -1 on malformed input
int64_t lz4_block(const uint8_t* in, size_t n, uint8_t* out, size_t cap) {
  size_t ip = 0, op = 0;
  for (;;) {
    if (ip >= n) return -1;
    const unsigned token = in[ip++];
    size_t lit = token >> 4;
    if (lit == 15) {
      unsigned b;
      do {
        if (ip >= n) return -1;
        b = in[ip++];
        lit += b;
      } while (b == 255);
    }
    if (lit > n - ip || lit > cap - op) return -1;
    std::memcpy(out + op, in + ip, lit);
    ip += lit;
    op += lit;
    if (ip == n) return static_cast<int64_t>(op);  // the last sequence: literals only
    if (n - ip < 2) return -1;
    const size_t off = size_t(in[ip]) | (size_t(in[ip + 1]) << 8);
    ip += 2;
    if (off == 0 || off > op) return -1;
    size_t len = (token & 15u) + 4;
    if ((token & 15u) == 15) {
      unsigned b;
      do {
        if (ip >= n) return -1;
        b = in[ip++];
        len += b;
      } while (b == 255);
    }
    if (len > cap - op) return -1;
    // byte by byte: the match may overlap what it writes (off < len)
    const uint8_t* src = out + op - off;
    for (size_t i = 0; i < len; ++i) out[op + i] = src[i];
    op += len;
  }
}

// 8x8 bit-matrix transpose: bit c of byte r <-> bit r of byte c
uint64_t transpose8(uint64_t x) {
  uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  return x ^ t ^ (t << 28);
}

// bit planes of n elements (n a multiple of 8) of E bytes -> the elements
void untranspose(const uint8_t* in, uint8_t* out, size_t n, size_t E) {
  const size_t nq = n / 8;
  for (size_t b = 0; b < E; ++b) {
    const uint8_t* pl = in + 8 * b * nq;  // planes 8b .. 8b+7
    for (size_t q = 0; q < nq; ++q) {
      uint64_t x = 0;
      for (int k = 0; k < 8; ++k) x |= uint64_t(pl[k * nq + q]) << (8 * k);
      x = transpose8(x);  // byte m: bit k = bit m of plane 8b+k's byte q = bit k of element 8q+m's byte b
      for (int m = 0; m < 8; ++m) out[(8 * q + m) * E + b] = static_cast<uint8_t>(x >> (8 * m));
    }
  }
}

}  // namespace
}  // namespace zdt

using namespace zdt;

static int bshuf_decompress(const void* in, size_t in_len, void* out, size_t out_len, size_t elem_size,
                            size_t block_size, int compression) {
  if ((in == nullptr && in_len > 0) || (out == nullptr && out_len > 0) || elem_size == 0)
    return fail(ZDT_EINVAL, "bitshuffle: bad buffers or element size");
  if (compression != 0 && compression != 2)
    return fail(ZDT_ENOTSUP, "bitshuffle: compression %d (only none = 0 and LZ4 = 2)", compression);
  if (out_len % elem_size != 0) return fail(ZDT_EINVAL, "bitshuffle: output size is not whole elements");
  const uint8_t* ip = static_cast<const uint8_t*>(in);
  uint8_t* op = static_cast<uint8_t*>(out);
  size_t left = in_len;
  if (compression == 2) {
    if (left < 12) return fail(ZDT_EINVAL, "bitshuffle: chunk shorter than its header");
    if (be64(ip) != out_len)
      return fail(ZDT_EINVAL, "bitshuffle: chunk holds %llu bytes, %zu expected", (unsigned long long)be64(ip),
                  out_len);
    const uint32_t bb = be32(ip + 8);
    if (bb == 0 || bb % elem_size != 0) return fail(ZDT_EINVAL, "bitshuffle: bad block size %u bytes", bb);
    block_size = bb / elem_size;
    ip += 12;
    left -= 12;
  } else {
    if (block_size == 0) {  // the library's default: 8 KiB of elements, a multiple of 8, at least 128
      block_size = (8192 / elem_size) / 8 * 8;
      if (block_size < 128) block_size = 128;
    }
    if (in_len != out_len) return fail(ZDT_EINVAL, "bitshuffle: uncompressed chunk of the wrong size");
  }
  if (block_size % 8 != 0) return fail(ZDT_EINVAL, "bitshuffle: block size %zu not a multiple of 8", block_size);
  // locate every block first (LZ4 blocks are size-prefixed), then decode
  // them on several threads: blocks are independent
  struct Blk {
    const uint8_t* src;
    size_t src_len;  // compressed bytes (LZ4), or the block's bytes
    size_t out_off, nb;
  };
  std::vector<Blk> blks;
  const size_t n = out_len / elem_size;
  size_t done = 0;  // elements
  while (n - done >= 8) {
    const size_t nb = (n - done >= block_size) ? block_size : (n - done) / 8 * 8;
    const size_t bytes = nb * elem_size;
    if (compression == 2) {
      if (left < 4) return fail(ZDT_EINVAL, "bitshuffle: truncated block header");
      const uint32_t cs = be32(ip);
      ip += 4;
      left -= 4;
      if (cs > left) return fail(ZDT_EINVAL, "bitshuffle: block of %u bytes past the chunk end", cs);
      blks.push_back({ip, cs, done * elem_size, nb});
      ip += cs;
      left -= cs;
    } else {
      if (bytes > left) return fail(ZDT_EINVAL, "bitshuffle: truncated chunk");
      blks.push_back({ip, bytes, done * elem_size, nb});
      ip += bytes;
      left -= bytes;
    }
    done += nb;
  }
  const size_t rest = (n - done) * elem_size;  // fewer than 8 elements, stored raw
  if (left != rest) return fail(ZDT_EINVAL, "bitshuffle: %zu trailing bytes, %zu expected", left, rest);
  std::atomic<int64_t> bad{-1};  // element offset of a malformed block (any one of them)
  std::atomic<bool> oom{false};
  auto work = [&](size_t b0, size_t b1) {
    std::vector<uint8_t> tmp;
    try {
      tmp.resize(compression == 2 ? block_size * elem_size : 0);
    } catch (...) {  // nothing may leave a thread
      oom.store(true);
      return;
    }
    for (size_t q = b0; q < b1; ++q) {
      const Blk& k = blks[q];
      const uint8_t* planes = k.src;
      if (compression == 2) {
        const size_t bytes = k.nb * elem_size;
        if (lz4_block(k.src, k.src_len, tmp.data(), bytes) != static_cast<int64_t>(bytes)) {
          int64_t expect = -1;
          bad.compare_exchange_strong(expect, static_cast<int64_t>(k.out_off / elem_size));
          return;
        }
        planes = tmp.data();
      }
      untranspose(planes, op + k.out_off, k.nb, elem_size);
    }
  };
  const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
  const size_t nth = std::min<size_t>({16, hw, (blks.size() + 7) / 8});
  bool serial = nth <= 1;
  if (!serial) {
    std::vector<std::thread> pool;
    const size_t per = (blks.size() + nth - 1) / nth;
    try {
      for (size_t i = 0; i < nth; ++i) {
        const size_t b0 = i * per, b1 = std::min(blks.size(), b0 + per);
        if (b0 < b1) pool.emplace_back(work, b0, b1);
      }
    } catch (...) {  // no thread to be had: the started ones finish, then one thread does it all
      serial = true;
    }
    for (auto& th : pool) th.join();
  }
  if (serial) {
    bad.store(-1);
    oom.store(false);
    work(0, blks.size());
  }
  if (oom.load()) return fail(ZDT_ENOMEM, "bitshuffle: out of host memory");
  if (bad.load() >= 0) return fail(ZDT_EINVAL, "bitshuffle: malformed LZ4 block at element %lld", (long long)bad.load());
  std::memcpy(op + done * elem_size, ip, rest);
  return ZDT_OK;
}

extern "C" int zdt_bshuf_decompress(const void* in, size_t in_len, void* out, size_t out_len, size_t elem_size,
                                    size_t block_size, int compression) {
  try {
    return bshuf_decompress(in, in_len, out, out_len, elem_size, block_size, compression);
  } catch (...) {  // the block list or a decode buffer could not be allocated
    return fail(ZDT_ENOMEM, "bitshuffle: out of host memory");
  }
}
