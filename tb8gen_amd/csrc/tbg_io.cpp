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
#include <array>
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
  // frames cover disjoint whole chunks: each offset is a multiple of kChunk,
  // seen once, and its frame holds exactly that chunk's bytes
  std::vector<uint8_t> seen((bytes + kChunk - 1) / kChunk, 0);
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
      if (h[1] > kChunk || h[1] == 0 || h[0] > bound || h[2] >= bytes || h[2] % kChunk ||
          h[1] != std::min<uint64_t>(kChunk, bytes - h[2]) || h[1] > left || seen[h[2] / kChunk]) {
        *err = std::string(path) + ": bad frame header";
        rc = -1;
        break;
      }
      seen[h[2] / kChunk] = 1;
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

// ----------------------------------------------------- Syzygy WDL files
//
// The layout src/probe.c maps (init_table 1210-1345, setup_pairs 1174-1208,
// setup_huffman 1061-1146, decompress_huff 1388-1456), written with the
// simplest compression it reads: every symbol a single WDL value (no
// Re-Pair pairs), a canonical Huffman code over the values present, 64-byte
// blocks of whole codewords (at most 65536 per block), and an index entry
// per 1024 positions.  The reader decodes the codes of length l as
// offset[l] + (code - base[l]), with base[] derived from the per-length
// offsets so that longer codes take the numerically smaller values.

namespace {

constexpr int kBlockLog = 6;   // 64-byte blocks
constexpr int kIdxBits = 10;   // an index entry per 1024 positions
constexpr uint64_t kBlockSyms = 65536;

void put16(std::vector<uint8_t> &o, uint32_t v)
{
  o.push_back((uint8_t)v);
  o.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t> &o, uint32_t v)
{
  put16(o, v & 0xffff);
  put16(o, v >> 16);
}

struct Packed {
  std::vector<uint8_t> head, index, sizes, data;
};

// Huffman code lengths of the values with nonzero count (deterministic:
// the two lightest nodes merge, ties to the lower node id)
void huffman_lengths(const uint64_t *cnt, int nv, int *len)
{
  struct Node { uint64_t w; int id, kids[2]; };
  std::vector<Node> nodes;
  std::vector<int> live;
  for (int v = 0; v < nv; v++) {
    len[v] = 0;
    if (cnt[v]) {
      nodes.push_back({ cnt[v], v, { -1, -1 } });
      live.push_back((int)nodes.size() - 1);
    }
  }
  int next_id = nv;
  while (live.size() > 1) {
    std::sort(live.begin(), live.end(), [&](int a, int b) {
      return nodes[a].w != nodes[b].w ? nodes[a].w < nodes[b].w : nodes[a].id < nodes[b].id;
    });
    Node n{ nodes[live[0]].w + nodes[live[1]].w, next_id++, { live[0], live[1] } };
    nodes.push_back(n);
    live.erase(live.begin(), live.begin() + 2);
    live.push_back((int)nodes.size() - 1);
  }
  // depth of every leaf
  std::vector<std::pair<int, int>> st{ { live[0], 0 } };
  while (!st.empty()) {
    auto [i, d] = st.back();
    st.pop_back();
    if (nodes[i].kids[0] < 0) len[nodes[i].id] = d;
    else
      for (int k : nodes[i].kids) st.push_back({ k, d + 1 });
  }
}

// One sub-table: the pairs header, index table, size table and data blocks
// of the values v[0 .. n) (0..4, or 0xff where any value will do).
int pack_side(uint8_t *v, uint64_t n, int threads, Packed *out, std::string *err)
{
  int nt = nthreads(threads);
  // counts of the values, then the free positions take the most frequent
  uint64_t cnt[5] = { 0 }, nfree = 0;
  {
    std::vector<std::array<uint64_t, 6>> part(nt);
    uint64_t per = (n + nt - 1) / nt;
    std::atomic<int> tid{ 0 };
    run_pool(nt, [&]() {
        int t = tid++;
        auto &c = part[t];
        c.fill(0);
        uint64_t a = std::min(n, per * t), b = std::min(n, a + per);
        for (uint64_t i = a; i < b; i++) c[v[i] < 5 ? v[i] : 5]++;
      });
    for (auto &c : part) {
      for (int k = 0; k < 5; k++) cnt[k] += c[k];
      nfree += c[5];
    }
  }
  int fill = 0;
  for (int k = 1; k < 5; k++)
    if (cnt[k] > cnt[fill]) fill = k;
  if (nfree) {
    uint64_t per = (n + nt - 1) / nt;
    std::atomic<int> tid{ 0 };
    run_pool(nt, [&]() {
        int t = tid++;
        uint64_t a = std::min(n, per * t), b = std::min(n, a + per);
        for (uint64_t i = a; i < b; i++)
          if (v[i] > 4) v[i] = (uint8_t)fill;
      });
    cnt[fill] += nfree;
  }
  int present = 0;
  for (int k = 0; k < 5; k++) present += cnt[k] != 0;
  std::vector<uint8_t> &h = out->head;
  if (present <= 1) {  // a constant sub-table: flag 0x80 and the value
    h.push_back(0x80);
    h.push_back((uint8_t)fill);
    return 0;
  }
  int len[5];
  huffman_lengths(cnt, 5, len);
  int minl = 64, maxl = 0;
  for (int k = 0; k < 5; k++)
    if (cnt[k]) {
      minl = std::min(minl, len[k]);
      maxl = std::max(maxl, len[k]);
    }
  int hl = maxl - minl + 1;
  // symbols: longest codes first, then by value
  std::vector<int> syms;
  for (int l = maxl; l >= minl; l--)
    for (int k = 0; k < 5; k++)
      if (cnt[k] && len[k] == l) syms.push_back(k);
  std::vector<uint32_t> nlen(hl, 0), offset(hl, 0), base(hl, 0);
  for (int k = 0; k < 5; k++)
    if (cnt[k]) nlen[len[k] - minl]++;
  for (int i = hl - 2; i >= 0; i--) offset[i] = offset[i + 1] + nlen[i + 1];
  for (int i = hl - 2; i >= 0; i--) base[i] = (base[i + 1] + nlen[i + 1]) / 2;
  if (base[0] + nlen[0] != (1u << minl)) {
    *err = "internal error: incomplete Huffman code";
    return -1;
  }
  uint32_t code[5] = { 0 };
  for (size_t s = 0; s < syms.size(); s++) {
    int k = syms[s], i = len[k] - minl;
    code[k] = base[i] + ((uint32_t)s - offset[i]);
  }
  // blocks: whole codewords, at most 512 bits and 65536 symbols each
  std::vector<uint64_t> start{ 0 };
  {
    uint32_t bits = 0;
    uint64_t c = 0;
    const uint32_t cap = 8u << kBlockLog;
    for (uint64_t i = 0; i < n; i++) {
      uint32_t l = (uint32_t)len[v[i]];
      if (bits + l > cap || c == kBlockSyms) {
        start.push_back(i);
        bits = 0;
        c = 0;
      }
      bits += l;
      c++;
    }
  }
  uint64_t nb = start.size();
  start.push_back(n);
  if (nb > 0xffffffffull) {
    *err = "table too large for 32-bit block numbers";
    return -1;
  }
  out->data.assign(nb << kBlockLog, 0);
  {
    std::atomic<uint64_t> next{ 0 };
    run_pool(nt, [&]() {
        for (uint64_t b0; (b0 = next.fetch_add(4096)) < nb;)
          for (uint64_t b = b0; b < std::min(nb, b0 + 4096); b++) {
            uint8_t *o = out->data.data() + (b << kBlockLog);
            uint32_t pos = 0;
            for (uint64_t i = start[b]; i < start[b + 1]; i++) {
              int k = v[i];
              for (int t = len[k] - 1; t >= 0; t--, pos++)
                if ((code[k] >> t) & 1) o[pos >> 3] |= (uint8_t)(0x80 >> (pos & 7));
            }
          }
      });
  }
  // size table: symbols per block minus one, plus one block past the end
  // so that index entries whose reference point lies beyond the last
  // position still name a block
  for (uint64_t b = 0; b < nb; b++) put16(out->sizes, (uint32_t)(start[b + 1] - start[b] - 1));
  put16(out->sizes, (uint32_t)(kBlockSyms - 1));
  // index: for entry m the block holding position m * 1024 + 512 and that
  // position's place in it
  uint64_t ni = (n + (1ull << kIdxBits) - 1) >> kIdxBits;
  for (uint64_t m = 0; m < ni; m++) {
    uint64_t P = (m << kIdxBits) + (1ull << (kIdxBits - 1));
    uint64_t b = nb;
    if (P < n) b = (uint64_t)(std::upper_bound(start.begin(), start.begin() + nb, P) - start.begin()) - 1;
    put32(out->index, (uint32_t)b);
    put16(out->index, (uint32_t)(P - start[b]));
  }
  h.push_back(0);  // flags: Huffman
  h.push_back((uint8_t)kBlockLog);
  h.push_back((uint8_t)kIdxBits);
  h.push_back(1);  // blocks in the size table beyond the data
  put32(h, (uint32_t)nb);
  h.push_back((uint8_t)maxl);
  h.push_back((uint8_t)minl);
  for (int i = 0; i < hl; i++) put16(h, offset[i]);
  put16(h, (uint32_t)syms.size());
  for (int k : syms) {  // a leaf symbol: value, then 0xfff as its second half
    h.push_back((uint8_t)k);
    h.push_back(0xf0);
    h.push_back(0xff);
  }
  if (syms.size() & 1) h.push_back(0);
  return 0;
}

}  // namespace

// Writes a WDL file: header (magic, flags, piece order per side), the
// sub-tables' pairs headers, index tables, size tables and 64-byte aligned
// data.  vals[t] (n[t] values) is sub-table t; nsides 2 for a split table.
int rtbw_write(const char *path, const uint8_t *pieces, int num, uint8_t order, int nsides,
               uint8_t *const *vals, const uint64_t *n, int threads, std::string *err)
{
  Packed pk[2];
  for (int t = 0; t < nsides; t++)
    if (pack_side(vals[t], n[t], threads, &pk[t], err)) return -1;
  std::vector<uint8_t> f;
  put32(f, 0x5d23e871u);  // WDL magic (src/probe.c:33)
  f.push_back(nsides == 2 ? 1 : 0);
  f.push_back((uint8_t)(order | (nsides == 2 ? order << 4 : 0)));
  for (int i = 0; i < num; i++) f.push_back((uint8_t)(pieces[i] | (nsides == 2 ? pieces[i] << 4 : 0)));
  if (f.size() & 1) f.push_back(0);
  for (int t = 0; t < nsides; t++) f.insert(f.end(), pk[t].head.begin(), pk[t].head.end());
  for (int t = 0; t < nsides; t++) f.insert(f.end(), pk[t].index.begin(), pk[t].index.end());
  for (int t = 0; t < nsides; t++) f.insert(f.end(), pk[t].sizes.begin(), pk[t].sizes.end());
  FILE *F = fopen(path, "wb");
  if (!F) {
    *err = std::string("cannot open ") + path + " for writing";
    return -1;
  }
  static const uint8_t zero[64] = { 0 };
  bool ok = fwrite(f.data(), 1, f.size(), F) == f.size();
  uint64_t at = f.size();
  for (int t = 0; t < nsides && ok; t++) {
    uint64_t pad = ((at + 63) & ~63ull) - at;  // data blocks start 64-byte aligned
    ok = fwrite(zero, 1, pad, F) == pad;
    ok = ok && fwrite(pk[t].data.data(), 1, pk[t].data.size(), F) == pk[t].data.size();
    at += pad + pk[t].data.size();
  }
  ok = ok && fwrite(zero, 1, 64, F) == 64;  // the decoder reads 8 bytes past a block
  ok = (fclose(F) == 0) && ok;
  if (!ok) {
    *err = std::string("write error on ") + path;
    return -1;
  }
  return 0;
}

}  // namespace tbg
