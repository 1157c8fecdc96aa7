// anneal_stream.hip -- sweep-mode annealing of large dense QUBOs (n > 1024).
//
// When R x N fp32 fields no longer fit the register file (N=8192, R=4096:
// 128 MiB), they live in HBM, in P copies (P = 2 by default).  Spins are
// decided in blocks of k = 64; block g of a segment covers spins
// 64*(g mod NB).. of sweep g / NB (NB = N/64), and its flips are kept as +/-
// bit masks in a ring of (P+1)m slots.  Two kinds of launch run side by side
// on two streams:
//  - decide (one work-group per 64 replicas): reads the 64x64 field tile of
//    block g from the last field copy, adds the flips of the blocks not yet
//    in that copy (bf16 hi + lo v_mfma_f32_32x32x16_bf16, fp32 accumulation),
//    decides the 64 spins in order (each replica's decisions split over a
//    lane quad, in-block corrections with the fp32 64x64 diagonal sub-block
//    of Q) and publishes the flip masks;
//  - pass (persistent, one work-group per CU by default): adds the flips of a
//    group of m blocks to every tile of the field array, from one copy into
//    the other -- the field array moves through HBM once per m blocks.
// Decide of block g in group H reads the copy after pass H-P, so pass H runs
// while groups H+1 .. H+P-1 are decided (the host side, capi_dense.hip,
// orders them with events).  Thresholds are the sweep-mode ones of the register-resident
// kernel: flip iff dE < neg_ln_u(w) * tscale[t], w = Philox word of counter
// (j>>2, replica, sweep, tag).
#include <algorithm>

#include "os_kernels.h"

namespace os {

typedef __bf16 sbf16x8 __attribute__((ext_vector_type(8)));
typedef float sf32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int kSB = 64;        // spins per block
constexpr int kST = kSB + 4;   // row stride of the transposed field tile (floats)

__device__ __forceinline__ sf32x16 smfma(uint4 a, uint4 b, sf32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(sbf16x8, a),
                                                 __builtin_bit_cast(sbf16x8, b), c, 0, 0, 0);
}

constexpr int kPRows = 256;   // pass launch: replicas per tile (4 waves x 64)
constexpr int kBStride = 272; // pass launch: bytes per staged operand column (256 + pad)
constexpr int kQ4 = 20;       // decide launch: floats per (spin, lane-in-quad) row of qd4

// 8 spins (bits sh..sh+7 of the +1 / -1 flip masks) as 8 bf16 +1 / -1 / 0:
// with y = a | a << 15, (y >> 2q) & 0x10001 puts bits 2q and 2q+1 of a at
// bits 0 and 16, and 0x3F80 / 0x8000 are bf16 1.0 and the sign bit.
__device__ __forceinline__ uint4 spins8(uint32_t plus, uint32_t minus, int sh) {
  const uint32_t a = ((plus | minus) >> sh) & 0xFFu, m = (minus >> sh) & 0xFFu;
  const uint32_t ya = a | (a << 15), ym = m | (m << 15);
  uint32_t v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t xa = (ya >> (2 * q)) & 0x10001u, xm = (ym >> (2 * q)) & 0x10001u;
    v[q] = xa * 0x3F80u + xm * 0x8000u;
  }
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// Operands of up to K blocks for one wave's 32x32 part of a 64x64 field
// tile: the flip masks (plus lo, plus hi, minus lo, minus hi) of its 32
// replicas and the bf16 hi / lo rows of Q in the B layout of
// v_mfma_f32_32x32x16_bf16.  qcol points at the wave's column of qpack for
// block 0 of its instance; the flip masks of block g are at
// dring[(g % ring) * T + drow].
template <int K>
struct TileOps {
  uint4 dd[K];
  uint4 B[K][8];
};

template <int K>
__device__ __forceinline__ void ops_load(TileOps<K> &o, int g0, int cnt, int np, int NB, int ring,
                                         size_t T, size_t drow, const uint4 *qcol,
                                         const uint4 *__restrict__ dring, int kh) {
#pragma unroll
  for (int t = 0; t < K; ++t) {
    if (t < cnt) {
      const int g = g0 + t;
      o.dd[t] = dring[(size_t)(g % ring) * T + drow];
      const uint4 *qb = qcol + (size_t)(g % NB) * np * 16;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        o.B[t][k] = qb[2 * k + kh];
        o.B[t][4 + k] = qb[8 + 2 * k + kh];
      }
    }
  }
}

// C += D_g * Qhi + D_g * Qlo for the (up to K) blocks in `o`, in order.
template <int K>
__device__ __forceinline__ void ops_apply(sf32x16 &C, const TileOps<K> &o, int cnt, int kh) {
#pragma unroll
  for (int t = 0; t < K; ++t) {
    if (t < cnt) {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const uint32_t pw = ks < 2 ? o.dd[t].x : o.dd[t].y, mw = ks < 2 ? o.dd[t].z : o.dd[t].w;
        const uint4 a = spins8(pw, mw, 16 * (ks & 1) + 8 * kh);
        C = smfma(a, o.B[t][ks], C);
        C = smfma(a, o.B[t][4 + ks], C);
      }
    }
  }
}

// Operands of up to K blocks for both 32-replica halves of one column half
// (shared-operand decide): the flip masks of replicas lc and 32 + lc and the
// Q rows of the wave's columns, loaded once for the two tiles.
template <int K>
struct TileOps2 {
  uint4 dd[K][2];
  uint4 B[K][8];
};

template <int K>
__device__ __forceinline__ void ops_load2(TileOps2<K> &o, int g0, int cnt, int NB, int ring,
                                          size_t T, size_t drow, const uint4 *qcol, int np,
                                          const uint4 *__restrict__ dring, int kh) {
#pragma unroll
  for (int t = 0; t < K; ++t) {
    if (t < cnt) {
      const int g = g0 + t;
      const uint4 *dr = dring + (size_t)(g % ring) * T + drow;
      o.dd[t][0] = dr[0];
      o.dd[t][1] = dr[32];
      const uint4 *qb = qcol + (size_t)(g % NB) * np * 16;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        o.B[t][k] = qb[2 * k + kh];
        o.B[t][4 + k] = qb[8 + 2 * k + kh];
      }
    }
  }
}

// C0 / C1 += D_g * Qhi + D_g * Qlo for the (up to K) blocks in `o`, in order:
// per tile the same sequence of accumulations as ops_apply.
template <int K>
__device__ __forceinline__ void ops_apply2(sf32x16 &C0, sf32x16 &C1, const TileOps2<K> &o,
                                           int cnt, int kh) {
#pragma unroll
  for (int t = 0; t < K; ++t) {
    if (t < cnt) {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        uint4 a[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const uint32_t pw = ks < 2 ? o.dd[t][h].x : o.dd[t][h].y;
          const uint32_t mw = ks < 2 ? o.dd[t][h].z : o.dd[t][h].w;
          a[h] = spins8(pw, mw, 16 * (ks & 1) + 8 * kh);
        }
        C0 = smfma(a[0], o.B[t][ks], C0);
        C1 = smfma(a[1], o.B[t][ks], C1);
        C0 = smfma(a[0], o.B[t][4 + ks], C0);
        C1 = smfma(a[1], o.B[t][4 + ks], C1);
      }
    }
  }
}

// A 32x32 part of a field tile in the C layout of v_mfma_f32_32x32x16_bf16:
// element i of lane l is row 8*(i/4) + 4*(l/32) + i%4, column l%32 (Fp is
// already offset by 4*(l/32) rows and l%32 columns).
__device__ __forceinline__ sf32x16 tile_load(const float *Fp, int np) {
  sf32x16 C;
#pragma unroll
  for (int i = 0; i < 16; ++i) C[i] = Fp[(size_t)(8 * (i >> 2) + (i & 3)) * np];
  return C;
}

__device__ __forceinline__ void tile_store(float *Fp, int np, const sf32x16 &C) {
#pragma unroll
  for (int i = 0; i < 16; ++i) Fp[(size_t)(8 * (i >> 2) + (i & 3)) * np] = C[i];
}

// Value of lane (lane & ~3) | ow of each quad (ow compile-time).
template <int ow>
__device__ __forceinline__ int quad_bcast(int v) {
  return __builtin_amdgcn_mov_dpp(v, ow * 0x55, 0xF, 0xF, false);
}

}  // namespace

// Pass launch: Fout = Fin + the flips of the cnt blocks g0...  Persistent
// grid over tiles of 256 replicas x 64 columns (column block fastest; with
// the grid size and np/64 multiples of 8, every tile of one column block is
// served by one XCD and its L2).  Wave w holds rows 64w.. of the tile as 2 x
// 2 accumulators of 32x32; the work-group stages the Q rows of K blocks for
// its 64 columns in LDS once and all four waves read them from there.
template <int K>
__global__ __launch_bounds__(256) void stream_pass_kernel(int np, int Rp, int I, int g0, int cnt,
                                                          int ring, size_t T,
                                                          const uint4 *__restrict__ qpack,
                                                          const float *__restrict__ Fin,
                                                          float *__restrict__ Fout,
                                                          const uint4 *__restrict__ dring) {
  __shared__ __align__(16) unsigned char bst[K][64 * kBStride];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int lc = lane & 31, kh = lane >> 5;
  const int NB = np / kSB, RG = (Rp + kPRows - 1) / kPRows;
  const int ntiles = NB * RG * I;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int cx = tile % NB, rg = (tile / NB) % RG, s = tile / (NB * RG);
    const int r64 = rg * 4 + w;  // the wave's 64-replica group
    const bool active = r64 * 64 < Rp;
    const size_t row0 = (size_t)s * Rp + (size_t)r64 * 64;
    sf32x16 acc[2][2];
    if (active) {
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          acc[rt][ct] =
              tile_load(Fin + (row0 + 32 * rt + 4 * kh) * np + cx * kSB + 32 * ct + lc, np);
    }
    for (int c0 = 0; c0 < cnt; c0 += K) {
      const int kc = min(K, cnt - c0);
      __syncthreads();  // the previous chunk's operands are consumed
#pragma unroll
      for (int t = 0; t < K; ++t) {
        if (t < kc) {
          const int g = g0 + c0 + t;
          const uint4 *src = qpack + ((size_t)(s * NB + g % NB) * np + cx * kSB) * 16;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int idx = tid + 256 * i, col = idx >> 4, part = idx & 15;
            *reinterpret_cast<uint4 *>(&bst[t][col * kBStride + part * 16]) = src[idx];
          }
        }
      }
      uint4 dd[K][2];
      if (active) {
#pragma unroll
        for (int t = 0; t < K; ++t)
          if (t < kc) {
            const int g = g0 + c0 + t;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
              dd[t][rt] = dring[(size_t)(g % ring) * T + row0 + 32 * rt + lc];
          }
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int t = 0; t < K; ++t) {
          if (t < kc) {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
              uint4 a[2];
#pragma unroll
              for (int rt = 0; rt < 2; ++rt) {
                const uint32_t pw = ks < 2 ? dd[t][rt].x : dd[t][rt].y;
                const uint32_t mw = ks < 2 ? dd[t][rt].z : dd[t][rt].w;
                a[rt] = spins8(pw, mw, 16 * (ks & 1) + 8 * kh);
              }
#pragma unroll
              for (int ct = 0; ct < 2; ++ct) {
                const unsigned char *bc = &bst[t][(32 * ct + lc) * kBStride];
                const uint4 bh = *reinterpret_cast<const uint4 *>(bc + (2 * ks + kh) * 16);
                const uint4 bl = *reinterpret_cast<const uint4 *>(bc + 128 + (2 * ks + kh) * 16);
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                  acc[rt][ct] = smfma(a[rt], bh, acc[rt][ct]);
                  acc[rt][ct] = smfma(a[rt], bl, acc[rt][ct]);
                }
              }
            }
          }
        }
      }
    }
    if (active) {
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          tile_store(Fout + (row0 + 32 * rt + 4 * kh) * np + cx * kSB + 32 * ct + lc, np,
                     acc[rt][ct]);
    }
  }
}

// Decide launch for block g (block b = g mod NB of sweep `sweep`): grid
// (Rp/64, I), 256 threads.  The work-group reads its 64x64 tile of block b
// from Fin (the fields after blocks .. gb-1) and adds the flips of blocks gb
// .. g-1 in registers; everything it reads is requested up front, so the
// decisions wait for one memory round trip per kChunk pending blocks.  The
// 64 decisions of a replica are split over a quad of lanes: wave w takes
// replicas 16w + lane/4, lane r4 = lane%4 owns spins j = r4 (mod 4) -- it
// decides them and keeps their in-block corrections c[j] -- and each
// decision reaches the quad by a DPP broadcast, after which every lane
// corrects its own later spins with the fp32 diagonal sub-block of Q.  The
// waves run at raised priority: the launch shares the CUs with the pass
// launch and is on the critical path.
//
// KC: pending blocks whose operands load in one round trip; PIPE: the next
// chunk's operands are requested before the current chunk is applied; SH:
// waves 0 and 1 each add the pending flips to both replica halves of their
// column half (one operand load for the two tiles), while waves 2 and 3 make
// all the thresholds.
template <int KC, bool PIPE, bool SH>
__global__ __launch_bounds__(256) void stream_decide_kernel(
    int n, int np, int Rp, int g, int gb, int ring, size_t T, int sweep, float ts, uint32_t k0,
    uint32_t k1, uint32_t inst_base, int last_of_sweep, const uint4 *__restrict__ qpack,
    const float *__restrict__ qdb, const float *__restrict__ Fin, uint32_t *__restrict__ xbits,
    uint4 *__restrict__ dring, double *__restrict__ energy, double *__restrict__ ebest,
    uint32_t *__restrict__ bestbits) {
  __shared__ __align__(16) float fl[kSB][kST];        // fields of the block, [spin][replica]
  __shared__ __align__(16) float qd4[kSB][4][kQ4];    // Q[j][4i + r4] at [j][r4][i]
  __shared__ float thr[kSB][kSB + 1];                 // thresholds [replica][spin]
  __shared__ uint32_t newx[kSB][2];                   // the block's new state words
  __shared__ int improved[kSB];                       // replica's best energy improved
  __builtin_amdgcn_s_setprio(3);
  const int ry = blockIdx.x, s = blockIdx.y;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int lc = lane & 31, kh = lane >> 5, mt = w >> 1, nt = w & 1;
  const int NB = np / kSB, npw = np / 32, b = g % NB, cnt = g - gb;
  const size_t row0 = (size_t)s * Rp + (size_t)ry * 64, drow = row0 + 32 * mt + lc;
  const uint32_t tag = ((inst_base + (uint32_t)s) << 8) | TAG_SWEEP_THR;
  // requests: diagonal sub-block, field tile, pending flips and Q rows, and
  // the decided replica's state words and energies
  const float4 *qsrc = reinterpret_cast<const float4 *>(qdb + ((size_t)s * NB + b) * kSB * kSB);
  float4 qv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) qv[i] = qsrc[tid + 256 * i];
  const size_t off = (row0 + 32 * mt + 4 * kh) * np + b * kSB + 32 * nt + lc;
  const uint4 *qcol = qpack + ((size_t)s * NB * np + b * kSB + 32 * nt + lc) * 16;
  const int rep = 16 * w + (lane >> 2), r4 = lane & 3;
  const size_t rr = row0 + rep;
  const uint2 xw0 = *reinterpret_cast<const uint2 *>(xbits + rr * npw + 2 * b);
  const double e_old = energy[rr];
  const double e_best = last_of_sweep ? ebest[rr] : 0.0;
  // thresholds: 64 replicas x 16 Philox calls (4 spins each); item it
  // (0 .. 1023) is replica it / 16, spins 4 (it % 16) ..
  auto threshold4 = [&](int it) {
    const int r = it >> 4, q4 = it & 15;
    const int j0 = b * kSB + 4 * q4;
    const uint4 ph = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), (uint32_t)(ry * 64 + r),
                                              (uint32_t)sweep, tag), k0, k1);
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const float x = neg_ln_u(pick(ph, jj)) * ts;
      thr[r][4 * q4 + jj] = j0 + jj < n ? x : -INFINITY;
    }
  };
  if (SH) {
    if (mt == 0) {
      // both replica halves of column half nt (offset 32 rows for the second)
      sf32x16 C0 = tile_load(Fin + off, np), C1 = tile_load(Fin + off + (size_t)32 * np, np);
      TileOps2<KC> o;
      for (int c0 = 0; c0 < cnt; c0 += KC) {
        ops_load2(o, gb + c0, cnt - c0, NB, ring, T, row0 + lc, qcol, np, dring, kh);
        ops_apply2(C0, C1, o, cnt - c0, kh);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        *reinterpret_cast<float4 *>(&fl[32 * nt + lc][8 * q + 4 * kh]) =
            make_float4(C0[4 * q], C0[4 * q + 1], C0[4 * q + 2], C0[4 * q + 3]);
        *reinterpret_cast<float4 *>(&fl[32 * nt + lc][32 + 8 * q + 4 * kh]) =
            make_float4(C1[4 * q], C1[4 * q + 1], C1[4 * q + 2], C1[4 * q + 3]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) threshold4(tid - 128 + 128 * i);
    }
  } else {
    sf32x16 C = tile_load(Fin + off, np);
    TileOps<KC> o;
    ops_load(o, gb, cnt, np, NB, ring, T, drow, qcol, dring, kh);
#pragma unroll
    for (int i = 0; i < 4; ++i) threshold4(tid + 256 * i);
    if (PIPE) {
      TileOps<KC> o2;
      for (int c0 = 0; c0 < cnt; c0 += KC) {
        if (c0 + KC < cnt) ops_load(o2, gb + c0 + KC, cnt - c0 - KC, np, NB, ring, T, drow, qcol,
                                    dring, kh);
        ops_apply(C, o, cnt - c0, kh);
        o = o2;
      }
    } else {
      ops_apply(C, o, cnt, kh);
      for (int c0 = KC; c0 < cnt; c0 += KC) {
        ops_load(o, gb + c0, cnt - c0, np, NB, ring, T, drow, qcol, dring, kh);
        ops_apply(C, o, cnt - c0, kh);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<float4 *>(&fl[32 * nt + lc][32 * mt + 8 * q + 4 * kh]) =
          make_float4(C[4 * q], C[4 * q + 1], C[4 * q + 2], C[4 * q + 3]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int idx = tid + 256 * i, j = idx >> 4, c4 = idx & 15;  // Q[j][4*c4 .. 4*c4+3]
    qd4[j][0][c4] = qv[i].x;
    qd4[j][1][c4] = qv[i].y;
    qd4[j][2][c4] = qv[i].z;
    qd4[j][3][c4] = qv[i].w;
  }
  __syncthreads();
  uint32_t x0 = xw0.x, x1 = xw0.y;
  // this lane's spins 4i + r4: fields (corrected in place as spins are
  // decided) and thresholds, in registers
  float f[kSB / 4], th[kSB / 4];
#pragma unroll
  for (int i = 0; i < kSB / 4; ++i) {
    f[i] = fl[4 * i + r4][rep];
    th[i] = thr[rep][4 * i + r4];
  }
  float desum = 0.0f;
  // row j of the diagonal sub-block, this lane's columns 4i + r4 (i >= q0*4
  // only), loaded one step ahead
  float4 qc[4], qn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) qc[q] = reinterpret_cast<const float4 *>(qd4[0][r4])[q];
#pragma clang loop unroll(full)
  for (int j = 0; j < kSB; ++j) {
    if (j + 1 < kSB) {
#pragma unroll
      for (int q = ((j + 1) >> 2) >> 2; q < 4; ++q)
        qn[q] = reinterpret_cast<const float4 *>(qd4[j + 1][r4])[q];
    }
    uint32_t &xwd = j < 32 ? x0 : x1;
    const int bit = j & 31;
    const bool xj = (xwd >> bit) & 1u;
    // owner lane (r4 == j % 4): field and decision, broadcast to the quad
    const float h = f[j >> 2];
    const float de = xj ? -h : h;
    const int acc_own = de < th[j >> 2] ? 1 : 0;
    int acc, de_bits;
    switch (j & 3) {
      case 0: acc = quad_bcast<0>(acc_own); de_bits = quad_bcast<0>(__float_as_int(de)); break;
      case 1: acc = quad_bcast<1>(acc_own); de_bits = quad_bcast<1>(__float_as_int(de)); break;
      case 2: acc = quad_bcast<2>(acc_own); de_bits = quad_bcast<2>(__float_as_int(de)); break;
      default: acc = quad_bcast<3>(acc_own); de_bits = quad_bcast<3>(__float_as_int(de)); break;
    }
    const float d = acc ? (xj ? -1.0f : 1.0f) : 0.0f;
    xwd ^= acc ? (1u << bit) : 0u;
    // every lane of the quad sums the accepted dE in spin order
    desum += acc ? __int_as_float(de_bits) : 0.0f;
    // this lane's later spins 4i + r4 (i >= j/4; the decided ones among them
    // are not read again, and Q's diagonal is zero)
#pragma unroll
    for (int q = (j >> 2) >> 2; q < 4; ++q) {
      if (4 * q >= (j >> 2)) f[4 * q] = fmaf(qc[q].x, d, f[4 * q]);
      if (4 * q + 1 >= (j >> 2)) f[4 * q + 1] = fmaf(qc[q].y, d, f[4 * q + 1]);
      if (4 * q + 2 >= (j >> 2)) f[4 * q + 2] = fmaf(qc[q].z, d, f[4 * q + 2]);
      if (4 * q + 3 >= (j >> 2)) f[4 * q + 3] = fmaf(qc[q].w, d, f[4 * q + 3]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) qc[q] = qn[q];
  }
  if (r4 == 0) {
    const uint32_t w0 = xw0.x, w1 = xw0.y;
    *reinterpret_cast<uint2 *>(xbits + rr * npw + 2 * b) = make_uint2(x0, x1);
    const uint32_t f0 = x0 ^ w0, f1 = x1 ^ w1;
    dring[(size_t)(g % ring) * T + rr] = make_uint4(f0 & ~w0, f1 & ~w1, f0 & w0, f1 & w1);
    const double e_now = e_old + (double)desum;
    energy[rr] = e_now;
    if (last_of_sweep) {
      const bool better = e_now < e_best;
      if (better) ebest[rr] = e_now;
      improved[rep] = better ? 1 : 0;
      newx[rep][0] = x0;
      newx[rep][1] = x1;
    }
  }
  if (!last_of_sweep) return;
  // end of a sweep: the improved replicas' states become their best states,
  // copied by the whole work-group (the block's words from LDS)
  __syncthreads();
  for (int k = tid; k < kSB * npw; k += 256) {
    const int r = k / npw, q = k - r * npw;
    if (!improved[r]) continue;
    const size_t row = row0 + r;
    uint32_t v;
    if (q == 2 * b) v = newx[r][0];
    else if (q == 2 * b + 1) v = newx[r][1];
    else v = xbits[row * npw + q];
    bestbits[row * npw + q] = v;
  }
}

// qpack[s][b][j][0..63] = hi(Q[64b+k][j]), [64..127] = lo(...), diagonal zeroed;
// qdb[s][b][k][c] = fp32 Q[64b+k][64b+c] (diagonal sub-blocks, diagonal zeroed)
__global__ __launch_bounds__(256) void pack_q64_kernel(int n, int np, const double *__restrict__ q,
                                                       uint16_t *__restrict__ qpack,
                                                       float *__restrict__ qdb) {
  const int s = blockIdx.y;
  const int NB = np / kSB;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;  // (b, j, k)
  if (t >= (size_t)NB * np * kSB) return;
  const int k = (int)(t & 63);
  const int j = (int)((t >> 6) % np);
  const int b = (int)((t >> 6) / np);
  const int i = kSB * b + k;
  const double *qs = q + (size_t)s * n * n;
  float qf = 0.0f;
  if (i < n && j < n && i != j) qf = (float)qs[(size_t)i * n + j];
  uint16_t hi = 0, lo = 0;
  if (qf != 0.0f) {
    uint32_t u = __float_as_uint(qf);
    u += 0x7fffu + ((u >> 16) & 1u);
    hi = (uint16_t)(u >> 16);
    uint32_t r = __float_as_uint(qf - __uint_as_float((uint32_t)hi << 16));
    r += 0x7fffu + ((r >> 16) & 1u);
    lo = (uint16_t)(r >> 16);
  }
  uint16_t *dst = qpack + (size_t)s * NB * np * 128 + ((size_t)b * np + j) * 128;
  dst[k] = hi;
  dst[64 + k] = lo;
  if (j / kSB == b) qdb[((size_t)s * NB + b) * kSB * kSB + k * kSB + (j % kSB)] = qf;
}

int stream_block() { return kSB; }

void launch_pack_q64(hipStream_t st, int I, int n, int np, const double *q, uint16_t *qpack,
                     float *qdb) {
  const size_t work = (size_t)(np / kSB) * np * kSB;
  pack_q64_kernel<<<dim3((unsigned)((work + 255) / 256), I), 256, 0, st>>>(n, np, q, qpack, qdb);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_stream_pass(hipStream_t st, int I, int np, int Rp, int g0, int cnt, int ring,
                        int chunk, int grid, const uint4 *qpack, const float *Fin, float *Fout,
                        const uint4 *dring) {
  const size_t T = (size_t)I * Rp;
  const int ntiles = (np / kSB) * ((Rp + kPRows - 1) / kPRows) * I;
  grid = std::max(1, std::min(grid, ntiles));
  if (chunk >= 4)
    stream_pass_kernel<4><<<grid, 256, 0, st>>>(np, Rp, I, g0, cnt, ring, T, qpack, Fin, Fout,
                                                dring);
  else if (chunk >= 2)
    stream_pass_kernel<2><<<grid, 256, 0, st>>>(np, Rp, I, g0, cnt, ring, T, qpack, Fin, Fout,
                                                dring);
  else
    stream_pass_kernel<1><<<grid, 256, 0, st>>>(np, Rp, I, g0, cnt, ring, T, qpack, Fin, Fout,
                                                dring);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_stream_decide(hipStream_t st, int I, int n, int np, int Rp, int g, int gb, int ring,
                          int sweep, float ts, uint64_t seed, uint32_t inst_base,
                          bool last_of_sweep, const uint4 *qpack, const float *qdb,
                          const float *Fin, uint32_t *xbits, uint4 *dring, double *energy,
                          double *ebest, uint32_t *bestbits, int variant) {
  const dim3 grid(Rp / 64, I);
  const size_t T = (size_t)I * Rp;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const int lo = last_of_sweep ? 1 : 0;
#define OS_DECIDE(KC, PIPE, SH)                                                             \
  stream_decide_kernel<KC, PIPE, SH><<<grid, 256, 0, st>>>(                                  \
      n, np, Rp, g, gb, ring, T, sweep, ts, k0, k1, inst_base, lo, qpack, qdb, Fin, xbits,   \
      dring, energy, ebest, bestbits)
  switch (variant) {
    case 21: OS_DECIDE(2, true, false); break;
    case 30: OS_DECIDE(3, false, false); break;
    case 31: OS_DECIDE(3, true, false); break;
    case 40: OS_DECIDE(4, false, false); break;
    case 50: OS_DECIDE(5, false, false); break;
    case 60: OS_DECIDE(2, false, true); break;
    case 61: OS_DECIDE(1, false, true); break;
    case 62: OS_DECIDE(3, false, true); break;
    default: OS_DECIDE(2, false, false); break;
  }
#undef OS_DECIDE
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
