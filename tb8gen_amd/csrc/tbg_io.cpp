// tbg_io.cpp -- slice files in the reference's on-disk format.
//
// A file holds the bytes of one slice bitmap (or capture-exit region) as a
// sequence of frames, each a header of three 64-bit words {compressed
// size, chunk size, offset of the chunk in the slice} followed by the zstd
// (level 6) compression of that chunk; chunks are at most 50 KiB and the
// frames may come in any order (src/util.c:179-343, src/Makefile:5).
// Compression and decompression run on a pool of host threads, a batch of
// frames at a time, so a 5.5 GB slice never needs more than one batch of
// compressed buffers.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

namespace tbg {

constexpr uint64_t kChunk = 50 * 1024;  // COPYSIZE of src/util.c:180
constexpr int kLevel = 6;               // ZSTD_LEVEL of src/Makefile:5
constexpr uint64_t kBatch = 1024;       // frames compressed per batch

namespace {

struct Frame {
  uint64_t csize, chunk, off;
  std::vector<uint8_t> data;
};

int nthreads(int t)
{
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 16u) : 4;
}

// body() on nt threads, or inline when one is enough
template <class F> void run_pool(int nt, F body)
{
  if (nt <= 1) {
    body();
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++) pool.emplace_back(body);
  for (auto &t : pool) t.join();
}

}  // namespace

// 0 on success, -1 with a message in err on failure
int io_write(const char *path, const void *src, uint64_t bytes, int threads, std::string *err)
{
  FILE *F = fopen(path, "wb");
  if (!F) {
    *err = std::string("cannot open ") + path + " for writing";
    return -1;
  }
  const uint8_t *p = static_cast<const uint8_t *>(src);
  uint64_t nframes = (bytes + kChunk - 1) / kChunk;
  size_t bound = ZSTD_compressBound(kChunk);
  std::vector<Frame> fr(std::min<uint64_t>(nframes, kBatch));
  for (auto &f : fr) f.data.resize(bound);
  int rc = 0;
  for (uint64_t b0 = 0; b0 < nframes && !rc; b0 += kBatch) {
    uint64_t nb = std::min<uint64_t>(kBatch, nframes - b0);
    std::atomic<uint64_t> next{ 0 };
    std::atomic<int> bad{ 0 };
    run_pool((int)std::min<uint64_t>(nthreads(threads), nb), [&]() {
        ZSTD_CCtx *cx = ZSTD_createCCtx();
        for (uint64_t i; (i = next++) < nb;) {
          Frame &f = fr[i];
          f.off = (b0 + i) * kChunk;
          f.chunk = std::min<uint64_t>(kChunk, bytes - f.off);
          size_t c = ZSTD_compressCCtx(cx, f.data.data(), bound, p + f.off, f.chunk, kLevel);
          if (ZSTD_isError(c)) bad = 1;
          f.csize = c;
        }
        ZSTD_freeCCtx(cx);
      });
    if (bad) {
      *err = "zstd compression failed";
      rc = -1;
      break;
    }
    for (uint64_t i = 0; i < nb && !rc; i++) {
      uint64_t h[3] = { fr[i].csize, fr[i].chunk, fr[i].off };
      if (fwrite(h, sizeof h, 1, F) != 1 || fwrite(fr[i].data.data(), 1, fr[i].csize, F) != fr[i].csize) {
        *err = std::string("write error on ") + path;
        rc = -1;
      }
    }
  }
  if (fclose(F) != 0 && !rc) {
    *err = std::string("write error on ") + path;
    rc = -1;
  }
  return rc;
}

int io_read(const char *path, void *dst, uint64_t bytes, int threads, std::string *err)
{
  FILE *F = fopen(path, "rb");
  if (!F) {
    *err = std::string("cannot open ") + path;
    return -1;
  }
  uint8_t *d = static_cast<uint8_t *>(dst);
  uint64_t left = bytes;
  int rc = 0;
  size_t bound = ZSTD_compressBound(kChunk);
  std::vector<Frame> fr;
  while (left && !rc) {
    // a batch of frames from the file, then decompressed in parallel
    fr.clear();
    while (left && fr.size() < kBatch) {
      uint64_t h[3];
      if (fread(h, sizeof h, 1, F) != 1) {
        *err = std::string(path) + ": truncated (frame header)";
        rc = -1;
        break;
      }
      if (h[1] > kChunk || h[1] == 0 || h[0] > bound || h[2] > bytes || h[1] > bytes - h[2] ||
          h[1] > left) {
        *err = std::string(path) + ": bad frame header";
        rc = -1;
        break;
      }
      Frame f;
      f.csize = h[0];
      f.chunk = h[1];
      f.off = h[2];
      f.data.resize(f.csize);
      if (fread(f.data.data(), 1, f.csize, F) != f.csize) {
        *err = std::string(path) + ": truncated (frame data)";
        rc = -1;
        break;
      }
      left -= f.chunk;
      fr.push_back(std::move(f));
    }
    if (rc) break;
    std::atomic<uint64_t> next{ 0 };
    std::atomic<int> bad{ 0 };
    run_pool((int)std::min<uint64_t>(nthreads(threads), fr.size()), [&]() {
        ZSTD_DCtx *dx = ZSTD_createDCtx();
        for (uint64_t i; (i = next++) < fr.size();) {
          Frame &f = fr[i];
          size_t r = ZSTD_decompressDCtx(dx, d + f.off, f.chunk, f.data.data(), f.csize);
          if (ZSTD_isError(r) || r != f.chunk) bad = 1;
        }
        ZSTD_freeDCtx(dx);
      });
    if (bad) {
      *err = std::string(path) + ": corrupt frame";
      rc = -1;
    }
  }
  fclose(F);
  return rc;
}

}  // namespace tbg
