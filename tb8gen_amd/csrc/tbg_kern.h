// tbg_kern.h -- kernel entry points and the capture-table reference shared
// by the kernel and host translation units.
#pragma once
#include "tbg_dev.h"

namespace tbg {

// A finished smaller-material table that capture exits of set k land in.
struct SubRef {
  const uint16_t *codes[2];  // [462 * size] codes per side to move
  Idx ii;                    // its index layout
  uint64_t map;              // packed: its slot j <- our slot map[j]
  int num;                   // its piece count
  int flip;                  // 1: colours swapped (probe.c:1554-1565)
  int present;               // 0: a bare pair of kings (always a draw)
};

__global__ void k_pred(const Mat *Mp, int stm, const uint64_t *src, uint64_t *cand,
                       const uint64_t *won, int s0, int ns, uint64_t cps);
__global__ void k_pred_blk(const Mat *Mp, int stm, const uint64_t *src, uint64_t *cand,
                           const uint64_t *won, int s0, int ns, uint64_t blk_pos, uint64_t nblk,
                           uint64_t blk_words);
constexpr int PB_THREADS = 512;
constexpr int PB_QUEUE_BYTES = 8 * 1024 * 2;
__global__ void k_uncap(const Mat *Mp, int stm, int k, const uint64_t *src0, const uint64_t *src1,
                        int legality, uint64_t *cand, const uint64_t *won, int s0, int ns,
                        uint64_t cps);
__global__ void k_verify(const Mat *Mp, int stm, int n, uint64_t *cand, const uint64_t *wins_o,
                         const uint64_t *sub_o, const uint64_t *won, uint64_t *newL,
                         uint16_t *codes, int s0, int ns, uint64_t cps, unsigned long long *cnt);
__global__ void k_wsweep(uint64_t *cand, uint64_t *wins, uint64_t *newW, uint16_t *codes,
                         uint64_t nw, uint64_t words, uint64_t size, uint64_t w0, int n,
                         unsigned long long *cnt);
__global__ void k_init(const Mat *Mp, uint64_t *winsW, uint64_t *winsB, uint64_t *mateW,
                       uint64_t *mateB, uint16_t *codesW, uint16_t *codesB, uint64_t *kexW0,
                       uint64_t *kexW1, uint64_t *kexL0, uint64_t *kexL1, int s0, int ns,
                       unsigned long long *cnt);
__global__ void k_resolve(const Mat *Mp, const SubRef *Rp, int c, int k, uint64_t *loss,
                          uint64_t *bloss, uint64_t *cwin, uint64_t *win, int s0, int ns);
__global__ void k_or(uint64_t *dst, const uint64_t *src, uint64_t nw);

// copies the shared tables to the device (defined with the kernels)
hipError_t upload_tabs(const Tabs *h);

}  // namespace tbg
