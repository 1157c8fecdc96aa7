// os_dense.h -- the dense batch behind the C ABI's os_dense handle.
#pragma once

#include <vector>

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
  // qpack is then [I][n/64][np][128], qdb [I][n/64][64][64], and `fields`
  // holds P copies of the field array (P the pipeline depth: each pass reads
  // one copy and writes the next)
  bool stream = false;
  DevBuf q64, qpack, qdb;
  DevBuf fields, xbits, bestbits, energy, ebest, tscale, scratch, drift, bytes, marg, status;
  DevBuf tsr;  // replica exchange: per-row temperature scales
  // population annealing: the gathered rows of states, fields and energies
  DevBuf pa_x, pa_f, pa_e, pa_anc;
  // replica exchange rounds on the GPU: slots, replicas at slots, scales
  DevBuf pt_slots;
  // stream path: flip masks of the last (P + 1) * group blocks (ring), and
  // the decide launches on `side`, beside the field passes on the context's
  // stream; ev_dec marks the last decide of a group, ev_pass[G % P] pass G
  DevBuf dblk;
  static constexpr int kGroup = 5;      // default blocks per field pass
  static constexpr int kPassChunk = 2;  // default operand blocks per round trip (pass)
  static constexpr int kPassWgsPerCu = 1;  // default persistent pass work-groups per CU (room for decides)
  static constexpr int kDecideVariant = 61;  // default decide operand chunking (launch_stream_decide)
  // pipeline depth P: the decides of group G read the fields after pass G - P
  // (so pass G - P + 1 .. G - 1 run beside them); 2 .. kMaxDepth.  Deeper
  // pipelines take the pass hand-offs off the decisions' path but give each
  // decide m more pending blocks, which cost more (P = 3 is 23% slower at
  // N=8192)
  static constexpr int kDepth = 2, kMaxDepth = 4;
  int group = kGroup;
  // streaming path: the autotuner's choice (0: not tuned, defaults apply)
  int tuned_group = 0, tuned_pass_wgs = 0;
  hipStream_t side = nullptr;
  hipEvent_t ev_seg = nullptr, ev_dec = nullptr, ev_pass[kMaxDepth] = {};
  std::vector<hipEvent_t> pass_ev;  // timing pairs around sampled passes
  ~Dense() {
    for (auto *b : {&q64, &qpack, &qdb, &fields, &xbits, &bestbits, &energy, &ebest, &tscale,
                    &scratch, &drift, &bytes, &marg, &status, &dblk, &tsr, &pa_x, &pa_f,
                    &pa_e, &pa_anc, &pt_slots})
      b->release();
    for (hipEvent_t e : {ev_seg, ev_dec})
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_pass)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : pass_ev) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
  }
};

// A one-instance dense problem of size n whose FP64 matrix `fill` writes on
// the context's stream (it receives the zeroed n x n device array); packs the
// MFMA operands and waits for the stream.  Defined in capi_dense.hip.
Dense *dense_create_device(Ctx &c, int n, void (*fill)(void *arg, double *q64), void *arg);

}  // namespace os
