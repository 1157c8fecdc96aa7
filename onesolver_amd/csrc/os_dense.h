// os_dense.h -- the dense batch behind the C ABI's os_dense handle.
#pragma once

#include "os_ctx.h"

namespace os {

// I same-size instances: Q as FP64 (q64 [I][n][n]), as bf16 hi+lo MFMA
// operands (qpack [I][n/16][np][32]) and as fp32 diagonal sub-blocks
// (qdb [I][n/16][16][16]); plus the per-run state of the last anneal.
struct Dense {
  Ctx *ctx = nullptr;
  int I = 1;  // instances (same n), stacked in every array
  int n = 0, np = 0, npw = 0;
  // n > dense_max_n(): fields in HBM, blocks of 64 spins (anneal_stream.hip);
  // qpack is then [I][n/64][np][128] and qdb [I][n/64][64][64]
  bool stream = false;
  DevBuf q64, qpack, qdb;
  DevBuf fields, xbits, bestbits, energy, ebest, tscale, scratch, drift, bytes, marg, status;
  DevBuf dblk;  // stream path: flip masks of the last two blocks
  ~Dense() {
    for (auto *b : {&q64, &qpack, &qdb, &fields, &xbits, &bestbits, &energy, &ebest, &tscale,
                    &scratch, &drift, &bytes, &marg, &status, &dblk})
      b->release();
  }
};

}  // namespace os
