// gemm_bf.hip -- projector GEMMs on the BF16 MFMA path.
//
// The operands are FP32 arrays; each value is split while it is staged into
// LDS into NS bfloat16 parts, x = x0 + x1 (+ x2), each part the
// round-to-nearest-even bfloat16 of the remainder of the previous ones (the
// subtractions are exact).  A product is the sum of the part products whose
// order stays within NS-1 (NS = 3: x0y0 x0y1 x1y0 x0y2 x1y1 x2y0; NS = 2:
// x0y0 x0y1 x1y0), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  With
// NS = 3 the dropped terms are below 2^-24 relative, the accuracy of an
// FP32 product; NS = 2 keeps about 16 bits.
//
// Same product shape as gemm.hip: D[p][q] = sum_r P(r,p) Q(r,q) on a
// 128 x 128 tile per block of 256 threads (2 x 2 waves of 64 x 64), or, when
// both output extents are multiples of 256 and NS = 2, a 256 x 256 tile per
// block of 512 threads (k_gemm_bfb), with
// P(r,p) = P[r + p*ldp] and Q(r,q) = Q[r + q*ldq] (TN) or Q[q + r*ldq] (NN).
// All operand extents are padded (rows to BK, columns to the tile edge, leading
// dimensions to multiples of 4) and the padding holds zeros, so loads need
// no bounds checks; stores are bounded by the valid extents.
//
// LDS holds the parts as [part][row][k] with k contiguous: the MFMA lane l
// reads row l%16 and the 8 consecutive k at 8*(l/16) as one 16-byte load
// (both operands use the same k order, so the products pair k correctly).
// The Q fragment is the MFMA's first operand, so accumulator e of lane l is
// D[p = l%16][q = 4(l/16) + e]: four consecutive q, stored as one vector.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "energy.h"

namespace qbg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int BT = 128;    // tile edge (p and q)
constexpr int BK = 32;     // reduction chunk (one MFMA k-step)
constexpr int LR = 40;     // LDS row pitch in bf16 (80 bytes: conflict-free 16-byte reads)

// round-to-nearest-even bfloat16 bits of x; `back` is the value as FP32
__device__ __forceinline__ unsigned short bf16_rne(float x, float& back) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  back = __uint_as_float(u);
  return (unsigned short)(u >> 16);
}

// split 8 values into NS parts of 8 bf16 each; the conversions are the
// hardware's round-to-nearest-even FP32 -> BF16 (packed, two per instruction)
template <int NS>
__device__ __forceinline__ void split8(const float* v, u16x8 (&out)[NS]) {
  float r[8];
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = v[e];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    bf16x8 h;
#pragma unroll
    for (int e = 0; e < 8; e++) h[e] = (__bf16)r[e];
    out[s] = __builtin_bit_cast(u16x8, h);
    if (s + 1 < NS) {
#pragma unroll
      for (int e = 0; e < 8; e++) r[e] -= (float)h[e];
    }
  }
}

// split 4 values into 2 parts of 4 bf16 each (as split8)
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void split4(const float* v, u16x4& hi, u16x4& lo) {
  bf16x4 h, l;
#pragma unroll
  for (int e = 0; e < 4; e++) h[e] = (__bf16)v[e];
#pragma unroll
  for (int e = 0; e < 4; e++) l[e] = (__bf16)(v[e] - (float)h[e]);
  hi = __builtin_bit_cast(u16x4, h);
  lo = __builtin_bit_cast(u16x4, l);
}

struct BfArgs {
  const float* P;
  long ldp;
  const float* Q;
  long ldq;
  int tp, tq;          // tiles along p and q
  int np, nq;          // valid output extent (stores)
  long nr;             // padded reduction extent (multiple of BK)
  long rper;           // reduction length of one split (multiple of BK)
  int pfast;           // block order: p tiles fastest (1) or q tiles fastest (0)
  void* out;           // out[split*sstride + q + p*ldo]
  long ldo, sstride;
  int accumulate;
  int vec;             // out and ldo allow the 4-wide vector stores
  // force epilogue (k_gemm_bfb with FE): D = Y(K, m) of the force product is
  // not stored; fpart[(m*3 + j)*ntq + q tile] = sum over the tile's G of
  // gx_j(G) Im(conj(anl(m, G)) Y(m, G))
  const float* fa;     // anl, fa[q + m*lda_fa]
  long lda_fa;
  const double* gx;    // gx[j*ngw + ig]
  int ngw, ntq;
  double* fpart;
  int nblk;            // tiles x splits; the grid is that rounded up to 8
};

// the hardware deals consecutive blocks of a grid out to the 8 XCDs round
// robin: renumber them so that each XCD runs one contiguous range of the
// nblk logical blocks (tiles sharing an operand slice stay on one L2); the
// grid is nblk rounded up to a multiple of 8 and the blocks past nblk return
__device__ __forceinline__ int xcd_block(int nblk) {
  const int b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  return b < nblk ? b : -1;
}

inline unsigned xcd_grid(BfArgs& g, int nblk) {
  g.nblk = nblk;
  return (unsigned)((nblk + 7) & ~7);
}

template <int NS, bool QR, typename TO>
__global__ __launch_bounds__(256) void k_gemm_bf(BfArgs g) {
  __shared__ __align__(16) unsigned short sP[NS][BT][LR];
  __shared__ __align__(16) unsigned short sQ[NS][BT][LR];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int b = xcd_block(g.nblk);
  if (b < 0) return;
  const int ntile = g.tp * g.tq;
  const int split = b / ntile, tt = b - split * ntile;
  int ip, iq;
  if (g.pfast) { ip = tt % g.tp; iq = tt / g.tp; }
  else { iq = tt % g.tq; ip = tt / g.tq; }
  const long p0 = (long)ip * BT, q0 = (long)iq * BT;
  const long rb = (long)split * g.rper, re = min(g.nr, rb + g.rper);
  // staging: r-contiguous operands as (row = 16 rows per 16 lanes, 16 k at
  // 16*((tid>>4)&1)): each 16-lane group writes 16 consecutive rows at one k
  // offset, the pattern of the fragment reads (no bank conflicts), and lanes
  // l, l+16 read the two halves of one 128-byte row segment;
  // q-contiguous Q as (q = tid&127, 16 k at 16*(tid>>7))
  const int srow = (tid & 15) | ((tid >> 5) << 4), sk = ((tid >> 4) & 1) * 16;
  const int qrow = QR ? srow : (tid & 127), qk = QR ? sk : (tid >> 7) * 16;
  const float* Pp = g.P + (p0 + srow) * g.ldp + sk;
  const float* Qp = QR ? g.Q + (q0 + srow) * g.ldq + sk : g.Q + q0 + qrow + (long)qk * g.ldq;
  float pv[16], qv[16];
  auto load = [&](long r0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float4 t = *reinterpret_cast<const float4*>(Pp + r0 + 4 * i);
      pv[4 * i] = t.x; pv[4 * i + 1] = t.y; pv[4 * i + 2] = t.z; pv[4 * i + 3] = t.w;
    }
    if (QR) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float4 t = *reinterpret_cast<const float4*>(Qp + r0 + 4 * i);
        qv[4 * i] = t.x; qv[4 * i + 1] = t.y; qv[4 * i + 2] = t.z; qv[4 * i + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) qv[i] = Qp[(r0 + i) * g.ldq];
    }
  };
  auto store = [&]() {
    u16x8 h[NS];
    split8<NS>(pv, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sP[s][srow][sk]) = h[s];
    split8<NS>(pv + 8, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sP[s][srow][sk + 8]) = h[s];
    split8<NS>(qv, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sQ[s][qrow][qk]) = h[s];
    split8<NS>(qv + 8, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sQ[s][qrow][qk + 8]) = h[s];
  };
  const int wp = (w & 1) * 64, wq = (w >> 1) * 64;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  load(rb);
  for (long r0 = rb; r0 < re; r0 += BK) {
    store();
    __syncthreads();
    if (r0 + BK < re) load(r0 + BK);
    bf16x8 a[NS][4], bq[NS][4];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[s][i] = *reinterpret_cast<const bf16x8*>(&sP[s][wp + 16 * i + fr][fk]);
        bq[s][i] = *reinterpret_cast<const bf16x8*>(&sQ[s][wq + 16 * i + fr][fk]);
      }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v4f c = acc[i][j];
        // smallest part products first
        if constexpr (NS == 3) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[2][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[1][j], a[1][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[2][j], a[0][i], c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[1][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[1][j], a[0][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[0][i], c, 0, 0, 0);
        acc[i][j] = c;
      }
    __syncthreads();
  }
  // lane: row p = .. + fr, four consecutive q from 4*(lane/16); whole
  // vectors when all four are inside the valid extent and the output is
  // aligned for them (g.vec: 16-byte base, ldo a multiple of 4; q is one)
  TO* o = (TO*)g.out + (long)split * g.sstride;
  const int fq = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long p = p0 + wp + 16 * i + fr;
      const long q = q0 + wq + 16 * j + fq;
      if (p >= g.np) continue;
      TO* d = o + q + p * g.ldo;
      if (g.vec && q + 3 < g.nq) {
        if constexpr (sizeof(TO) == 4) {
          float4 v = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
          if (g.accumulate) {
            const float4 u = *reinterpret_cast<const float4*>(d);
            v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
          }
          *reinterpret_cast<float4*>(d) = v;
        } else {
          double2 v0 = {(double)acc[i][j][0], (double)acc[i][j][1]};
          double2 v1 = {(double)acc[i][j][2], (double)acc[i][j][3]};
          if (g.accumulate) {
            const double2 u0 = *reinterpret_cast<const double2*>(d);
            const double2 u1 = *reinterpret_cast<const double2*>(d + 2);
            v0.x = u0.x + v0.x; v0.y = u0.y + v0.y;
            v1.x = u1.x + v1.x; v1.y = u1.y + v1.y;
          }
          *reinterpret_cast<double2*>(d) = v0;
          *reinterpret_cast<double2*>(d + 2) = v1;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (q + e < g.nq) {
            const TO v = (TO)acc[i][j][e];
            d[e] = g.accumulate ? (TO)(d[e] + v) : v;
          }
      }
    }
}

// 256 x 256 tiles for large products: 512 threads = 8 waves of 128 (p) x 64
// (q), so each staged operand row serves twice as many outputs as in the
// 128 x 128 kernel (half the operand traffic from L2 and the LDS writes per
// MFMA); fragments of P are read one 16-row group at a time to keep the
// registers of 2 waves per SIMD
constexpr int BB = 256;
template <int NS, bool QR, typename TO, bool FE = false>
__global__ __launch_bounds__(512) void k_gemm_bfb(BfArgs g) {
  __shared__ __align__(16) unsigned short sP[NS][BB][LR];
  __shared__ __align__(16) unsigned short sQ[NS][BB][LR];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int b = xcd_block(g.nblk);
  if (b < 0) return;
  const int ntile = g.tp * g.tq;
  const int split = b / ntile, tt = b - split * ntile;
  int ip, iq;
  if (g.pfast) { ip = tt % g.tp; iq = tt / g.tp; }
  else { iq = tt % g.tq; ip = tt / g.tq; }
  const long p0 = (long)ip * BB, q0 = (long)iq * BB;
  const long rb = (long)split * g.rper, re = min(g.nr, rb + g.rper);
  const int srow = (tid & 15) | ((tid >> 5) << 4), sk = ((tid >> 4) & 1) * 16;
  const int qrow = QR ? srow : (tid & 255), qk = QR ? sk : (tid >> 8) * 16;
  const float* Pp = g.P + (p0 + srow) * g.ldp + sk;
  const float* Qp = QR ? g.Q + (q0 + srow) * g.ldq + sk : g.Q + q0 + qrow + (long)qk * g.ldq;
  float pv[16], qv[16];
  auto load = [&](long r0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float4 t = *reinterpret_cast<const float4*>(Pp + r0 + 4 * i);
      pv[4 * i] = t.x; pv[4 * i + 1] = t.y; pv[4 * i + 2] = t.z; pv[4 * i + 3] = t.w;
    }
    if (QR) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float4 t = *reinterpret_cast<const float4*>(Qp + r0 + 4 * i);
        qv[4 * i] = t.x; qv[4 * i + 1] = t.y; qv[4 * i + 2] = t.z; qv[4 * i + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) qv[i] = Qp[(r0 + i) * g.ldq];
    }
  };
  auto store = [&]() {
    u16x8 h[NS];
    split8<NS>(pv, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sP[s][srow][sk]) = h[s];
    split8<NS>(pv + 8, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sP[s][srow][sk + 8]) = h[s];
    split8<NS>(qv, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sQ[s][qrow][qk]) = h[s];
    split8<NS>(qv + 8, h);
#pragma unroll
    for (int s = 0; s < NS; s++) *reinterpret_cast<u16x8*>(&sQ[s][qrow][qk + 8]) = h[s];
  };
  const int wp = (w & 1) * 128, wq = (w >> 1) * 64;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  v4f acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  load(rb);
  for (long r0 = rb; r0 < re; r0 += BK) {
    store();
    __syncthreads();
    if (r0 + BK < re) load(r0 + BK);
    bf16x8 bq[NS][4];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        bq[s][j] = *reinterpret_cast<const bf16x8*>(&sQ[s][wq + 16 * j + fr][fk]);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      bf16x8 a[NS];
#pragma unroll
      for (int s = 0; s < NS; s++)
        a[s] = *reinterpret_cast<const bf16x8*>(&sP[s][wp + 16 * i + fr][fk]);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v4f c = acc[i][j];
        if constexpr (NS == 3) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[2], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[1][j], a[1], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[2][j], a[0], c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[1][j], a[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[0], c, 0, 0, 0);
        acc[i][j] = c;
      }
    }
    __syncthreads();
  }
  if constexpr (FE) {
    // one projector fragment i at a time (3 double sums live): lane p =
    // p0 + wp + 16 i + fr, G vectors ig, ig+1 of each q fragment (q even:
    // re/im pairs); then the four lane groups sharing p, through LDS the
    // four waves sharing wp
    const int fq = 4 * (lane >> 4);
    const int wg = w >> 1;
    __syncthreads();                      // LDS tiles are free now
    double* red = reinterpret_cast<double*>(&sP[0][0][0]);   // [4][256][3]
    // the tile's G vectors q0/2 .. q0/2 + 127, staged once: gxs[c][ig - q0/2]
    double* gxs = red + 4 * BB * 3;
    for (int t = tid; t < 3 * (BB / 2); t += 512) {
      const int c = t / (BB / 2), k = t - c * (BB / 2);
      const long ig = (q0 >> 1) + k;
      gxs[t] = ig < g.ngw ? g.gx[(long)c * g.ngw + ig] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const long p = p0 + wp + 16 * i + fr;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long q = q0 + wq + 16 * j + fq;
        const int k = (int)((q - q0) >> 1);   // local G index (zero beyond ngw)
        const float4 a = *reinterpret_cast<const float4*>(g.fa + q + p * g.lda_fa);
        const double t0 = (double)a.y * acc[i][j][0] - (double)a.x * acc[i][j][1];
        const double t1 = (double)a.w * acc[i][j][2] - (double)a.z * acc[i][j][3];
        v0 += gxs[k] * t0 + gxs[k + 1] * t1;
        v1 += gxs[BB / 2 + k] * t0 + gxs[BB / 2 + k + 1] * t1;
        v2 += gxs[BB + k] * t0 + gxs[BB + k + 1] * t1;
      }
      v0 += __shfl_xor(v0, 16); v0 += __shfl_xor(v0, 32);
      v1 += __shfl_xor(v1, 16); v1 += __shfl_xor(v1, 32);
      v2 += __shfl_xor(v2, 16); v2 += __shfl_xor(v2, 32);
      if (lane < 16) {
        double* r = red + (wg * BB + wp + 16 * i + lane) * 3;
        r[0] = v0; r[1] = v1; r[2] = v2;
      }
    }
    __syncthreads();
    for (int t = tid; t < BB * 3; t += 512) {
      const int pl = t / 3, c = t - 3 * pl;
      const long p = p0 + pl;
      if (p < g.np) {
        const double sum = ((red[(0 * BB + pl) * 3 + c] + red[(1 * BB + pl) * 3 + c]) +
                            red[(2 * BB + pl) * 3 + c]) + red[(3 * BB + pl) * 3 + c];
        g.fpart[(p * 3 + c) * g.ntq + iq] = sum;
      }
    }
    return;
  }
  TO* o = (TO*)g.out + (long)split * g.sstride;
  const int fq = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long p = p0 + wp + 16 * i + fr;
      const long q = q0 + wq + 16 * j + fq;
      if (p >= g.np) continue;
      TO* d = o + q + p * g.ldo;
      if (g.vec && q + 3 < g.nq) {
        if constexpr (sizeof(TO) == 4) {
          float4 v = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
          if (g.accumulate) {
            const float4 u = *reinterpret_cast<const float4*>(d);
            v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
          }
          *reinterpret_cast<float4*>(d) = v;
        } else {
          double2 v0 = {(double)acc[i][j][0], (double)acc[i][j][1]};
          double2 v1 = {(double)acc[i][j][2], (double)acc[i][j][3]};
          if (g.accumulate) {
            const double2 u0 = *reinterpret_cast<const double2*>(d);
            const double2 u1 = *reinterpret_cast<const double2*>(d + 2);
            v0.x = u0.x + v0.x; v0.y = u0.y + v0.y;
            v1.x = u1.x + v1.x; v1.y = u1.y + v1.y;
          }
          *reinterpret_cast<double2*>(d) = v0;
          *reinterpret_cast<double2*>(d + 2) = v1;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (q + e < g.nq) {
            const TO v = (TO)acc[i][j][e];
            d[e] = g.accumulate ? (TO)(d[e] + v) : v;
          }
      }
    }
}

// 256 x 256 tiles with the 2-part split and two LDS stages (160 KB): the
// loop computes step i from one stage while the split parts of step i+1,
// loaded into registers during step i-1, go into the other stage and the
// loads of step i+2 are issued; one barrier per step.  An r-contiguous
// operand is loaded with eight consecutive lanes on the 128 contiguous bytes
// of one row (4 k per lane, rows 64 apart per load), so each load touches 8
// whole cache lines; Q along q (NN) as pairs of adjacent q (one 8-byte load
// per lane and k row).  Otherwise as k_gemm_bfb (same fragments, MFMA
// order, epilogues).
template <bool QR, typename TO, bool FE = false>
__global__ __launch_bounds__(512) void k_gemm_bfd(BfArgs g) {
  __shared__ __align__(16) unsigned short sm[2][2][2][BB][LR];   // [stage][P/Q][part][row][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int b = xcd_block(g.nblk);
  if (b < 0) return;
  const int ntile = g.tp * g.tq;
  const int split = b / ntile, tt = b - split * ntile;
  int ip, iq;
  if (g.pfast) { ip = tt % g.tp; iq = tt / g.tp; }
  else { iq = tt % g.tq; ip = tt / g.tq; }
  const long p0 = (long)ip * BB, q0 = (long)iq * BB;
  const long rb = (long)split * g.rper, re = min(g.nr, rb + g.rper);
  const long nsteps = (re - rb) / BK;
  // P (and Q when r-contiguous): rows lrow + 64 j (j < 4), 4 k at lk
  const int lrow = tid >> 3, lk = (tid & 7) * 4;
  // Q along q: rows qq, qq + 1 (adjacent q), 8 k at qk
  const int qq = 2 * (tid & 127), qk = (tid >> 7) * 8;
  const float* Pp = g.P + (p0 + lrow) * g.ldp + lk;
  const float* Qp = QR ? g.Q + (q0 + lrow) * g.ldq + lk : g.Q + q0 + qq + (long)qk * g.ldq;
  const long p64 = 64 * g.ldp, q64 = 64 * g.ldq;
  float pv[16], qv[16];
  auto load = [&](long r0) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float4 t = *reinterpret_cast<const float4*>(Pp + j * p64 + r0);
      pv[4 * j] = t.x; pv[4 * j + 1] = t.y; pv[4 * j + 2] = t.z; pv[4 * j + 3] = t.w;
    }
    if (QR) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float4 t = *reinterpret_cast<const float4*>(Qp + j * q64 + r0);
        qv[4 * j] = t.x; qv[4 * j + 1] = t.y; qv[4 * j + 2] = t.z; qv[4 * j + 3] = t.w;
      }
    } else {
      // qv[i] = Q(r0 + qk + i, qq), qv[8 + i] = Q(r0 + qk + i, qq + 1)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const float2 t = *reinterpret_cast<const float2*>(Qp + (r0 + i) * g.ldq);
        qv[i] = t.x; qv[8 + i] = t.y;
      }
    }
  };
  auto store = [&](int st) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      u16x4 hi, lo;
      split4(pv + 4 * j, hi, lo);
      *reinterpret_cast<u16x4*>(&sm[st][0][0][lrow + 64 * j][lk]) = hi;
      *reinterpret_cast<u16x4*>(&sm[st][0][1][lrow + 64 * j][lk]) = lo;
    }
    if (QR) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        u16x4 hi, lo;
        split4(qv + 4 * j, hi, lo);
        *reinterpret_cast<u16x4*>(&sm[st][1][0][lrow + 64 * j][lk]) = hi;
        *reinterpret_cast<u16x4*>(&sm[st][1][1][lrow + 64 * j][lk]) = lo;
      }
    } else {
      u16x8 h[2];
      split8<2>(qv, h);
      *reinterpret_cast<u16x8*>(&sm[st][1][0][qq][qk]) = h[0];
      *reinterpret_cast<u16x8*>(&sm[st][1][1][qq][qk]) = h[1];
      split8<2>(qv + 8, h);
      *reinterpret_cast<u16x8*>(&sm[st][1][0][qq + 1][qk]) = h[0];
      *reinterpret_cast<u16x8*>(&sm[st][1][1][qq + 1][qk]) = h[1];
    }
  };
  const int wp = (w & 1) * 128, wq = (w >> 1) * 64;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  v4f acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  if (nsteps > 0) {
    load(rb);
    store(0);
    if (nsteps > 1) load(rb + BK);
  }
  __syncthreads();
  for (long it = 0; it < nsteps; it++) {
    const int cur = (int)(it & 1);
    bf16x8 bq[2][4];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        bq[s][j] = *reinterpret_cast<const bf16x8*>(&sm[cur][1][s][wq + 16 * j + fr][fk]);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      bf16x8 a[2];
#pragma unroll
      for (int s = 0; s < 2; s++)
        a[s] = *reinterpret_cast<const bf16x8*>(&sm[cur][0][s][wp + 16 * i + fr][fk]);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v4f c = acc[i][j];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[1][j], a[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[0][j], a[0], c, 0, 0, 0);
        acc[i][j] = c;
      }
    }
    if (it + 1 < nsteps) {
      store(cur ^ 1);
      if (it + 2 < nsteps) load(rb + (it + 2) * BK);
    }
    __syncthreads();
  }
  if constexpr (FE) {
    const int fq = 4 * (lane >> 4);
    const int wg = w >> 1;
    double* red = reinterpret_cast<double*>(&sm[0][0][0][0][0]);   // [4][256][3]
    double* gxs = red + 4 * BB * 3;
    for (int t = tid; t < 3 * (BB / 2); t += 512) {
      const int c = t / (BB / 2), k = t - c * (BB / 2);
      const long ig = (q0 >> 1) + k;
      gxs[t] = ig < g.ngw ? g.gx[(long)c * g.ngw + ig] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const long p = p0 + wp + 16 * i + fr;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long q = q0 + wq + 16 * j + fq;
        const int k = (int)((q - q0) >> 1);
        const float4 a = *reinterpret_cast<const float4*>(g.fa + q + p * g.lda_fa);
        const double t0 = (double)a.y * acc[i][j][0] - (double)a.x * acc[i][j][1];
        const double t1 = (double)a.w * acc[i][j][2] - (double)a.z * acc[i][j][3];
        v0 += gxs[k] * t0 + gxs[k + 1] * t1;
        v1 += gxs[BB / 2 + k] * t0 + gxs[BB / 2 + k + 1] * t1;
        v2 += gxs[BB + k] * t0 + gxs[BB + k + 1] * t1;
      }
      v0 += __shfl_xor(v0, 16); v0 += __shfl_xor(v0, 32);
      v1 += __shfl_xor(v1, 16); v1 += __shfl_xor(v1, 32);
      v2 += __shfl_xor(v2, 16); v2 += __shfl_xor(v2, 32);
      if (lane < 16) {
        double* r = red + (wg * BB + wp + 16 * i + lane) * 3;
        r[0] = v0; r[1] = v1; r[2] = v2;
      }
    }
    __syncthreads();
    for (int t = tid; t < BB * 3; t += 512) {
      const int pl = t / 3, c = t - 3 * pl;
      const long p = p0 + pl;
      if (p < g.np) {
        const double sum = ((red[(0 * BB + pl) * 3 + c] + red[(1 * BB + pl) * 3 + c]) +
                            red[(2 * BB + pl) * 3 + c]) + red[(3 * BB + pl) * 3 + c];
        g.fpart[(p * 3 + c) * g.ntq + iq] = sum;
      }
    }
    return;
  }
  TO* o = (TO*)g.out + (long)split * g.sstride;
  const int fq = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long p = p0 + wp + 16 * i + fr;
      const long q = q0 + wq + 16 * j + fq;
      if (p >= g.np) continue;
      TO* d = o + q + p * g.ldo;
      if (g.vec && q + 3 < g.nq) {
        if constexpr (sizeof(TO) == 4) {
          float4 v = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
          if (g.accumulate) {
            const float4 u = *reinterpret_cast<const float4*>(d);
            v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
          }
          *reinterpret_cast<float4*>(d) = v;
        } else {
          double2 v0 = {(double)acc[i][j][0], (double)acc[i][j][1]};
          double2 v1 = {(double)acc[i][j][2], (double)acc[i][j][3]};
          if (g.accumulate) {
            const double2 u0 = *reinterpret_cast<const double2*>(d);
            const double2 u1 = *reinterpret_cast<const double2*>(d + 2);
            v0.x = u0.x + v0.x; v0.y = u0.y + v0.y;
            v1.x = u1.x + v1.x; v1.y = u1.y + v1.y;
          }
          *reinterpret_cast<double2*>(d) = v0;
          *reinterpret_cast<double2*>(d + 2) = v1;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (q + e < g.nq) {
            const TO v = (TO)acc[i][j][e];
            d[e] = g.accumulate ? (TO)(d[e] + v) : v;
          }
      }
    }
}

// out[q + p*ldo] = alpha sum_s partial[s][q + p*pitch] + gamma Q[q*ldq] P[p*ldp]
__global__ __launch_bounds__(256) void k_bf_reduce(const float* __restrict__ partial, int S,
                                                   int nq, int np, long pitch, long sstride,
                                                   double alpha,
                                                   double gamma, const float* __restrict__ Q,
                                                   long ldq, const double* __restrict__ P,
                                                   long ldp, double* __restrict__ out, long ldo) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nq * np) return;
  const int q = (int)(t % nq), p = (int)(t / nq);
  double s = 0.0;
  for (int k = 0; k < S; k++) s += (double)partial[(long)k * sstride + q + (long)p * pitch];
  double v = alpha * s;
  if (gamma != 0.0) v += gamma * (double)Q[(long)q * ldq] * P[(long)p * ldp];
  out[q + (long)p * ldo] = v;
}

// blocks of k_gemm_bf<NS, true, float> the current device runs at once
// (CUs x resident blocks per CU), cached per device
template <int NS>
int tn_slots() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  if (!cache[dev]) {
    int ncu = 0, occ = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, reinterpret_cast<const void*>(&k_gemm_bf<NS, true, float>), 256, 0) != hipSuccess)
      occ = 2;
    cache[dev] = std::max(1, ncu) * std::max(1, occ);
  }
  return cache[dev];
}

// blocks of k_gemm_bfb<2, true, float> the current device runs at once
int tnb_slots() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int ncu = 0, occ = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, reinterpret_cast<const void*>(&k_gemm_bfd<true, float>), 512, 0) != hipSuccess)
      occ = 1;
    cache[dev] = std::max(1, ncu) * std::max(1, occ);
  }
  return cache[dev];
}

template <int NS, bool QR, typename TO>
void launchb(BfArgs g, int nsplit, hipStream_t s) {
  static_assert(NS == 2, "256-edge tiles run the 2-part split only");
  const unsigned grid = xcd_grid(g, g.tp * g.tq * nsplit);
  hipLaunchKernelGGL((k_gemm_bfd<QR, TO>), dim3(grid), dim3(512), 0, s, g);
}

template <int NS, bool QR, typename TO>
void launch(BfArgs g, int nsplit, hipStream_t s) {
  const unsigned grid = xcd_grid(g, g.tp * g.tq * nsplit);
  hipLaunchKernelGGL((k_gemm_bf<NS, QR, TO>), dim3(grid), dim3(256), 0, s, g);
}

}  // namespace

int gemm_bf_tn_default_split(int Mp, int Np, long Kp, int parts, int edge) {
  const bool big = parts == 2 && edge == BB && Mp % BB == 0 && Np % BB == 0;
  const int bt = big ? BB : BT;
  const int ntile = (Np / bt) * (Mp / bt);
  const long ksteps = Kp / BK;
  // split the reduction so the blocks fill the device exactly once (a
  // second, partial round would run at a fraction of the occupancy)
  const int slots = big ? tnb_slots() : parts == 3 ? tn_slots<3>() : tn_slots<2>();
  long S = std::max<long>(1, slots / ntile);
  S = std::min<long>(S, std::max<long>(1, ksteps / 8));
  return (int)S;
}

int gemm_bf_tn(const float* A, long lda, const float* B, long ldb, int Mp, int Np, long Kp,
               int M, int N, const double* B64, long ldb64, double alpha, double gamma,
               double* out, long ldo, float* work, size_t work_floats, int parts,
               hipStream_t s, const BfTnPlan* plan) {
  if (Mp % BT || Np % BT || Kp % BK || lda % 4 || ldb % 4 || M > Mp || N > Np) return -1;
  BfArgs g{};
  g.P = B; g.ldp = ldb; g.Q = A; g.ldq = lda;
  const int edge = plan ? plan->edge : BB;
  const bool big = parts == 2 && edge == BB && Mp % BB == 0 && Np % BB == 0;
  const int bt = big ? BB : BT;
  g.tp = Np / bt; g.tq = Mp / bt;
  g.np = Np; g.nq = Mp;
  g.pfast = 0;
  const long ksteps = Kp / BK;
  long S = (plan && plan->split > 0) ? plan->split
                                      : gemm_bf_tn_default_split(Mp, Np, Kp, parts, edge);
  S = std::min<long>(S, ksteps);
  while (S > 1 && (size_t)S * Mp * Np > work_floats) S--;
  if ((size_t)S * Mp * Np > work_floats) return -1;
  g.nr = Kp;
  g.rper = ((ksteps + S - 1) / S) * BK;
  S = (Kp + g.rper - 1) / g.rper;
  g.out = work; g.ldo = Mp; g.sstride = (long)Mp * Np; g.accumulate = 0;
  g.vec = ((uintptr_t)work % 16 == 0) ? 1 : 0;
  if (big) launchb<2, true, float>(g, (int)S, s);
  else if (parts == 3) launch<3, true, float>(g, (int)S, s);
  else launch<2, true, float>(g, (int)S, s);
  const long tot = (long)M * N;
  hipLaunchKernelGGL(k_bf_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, work,
                     (int)S, M, N, (long)Mp, (long)Mp * Np, alpha, gamma, A, lda, B64, ldb64,
                     out, ldo);
  return 0;
}

template <typename TO>
static int bf_nn(const float* A, long lda, const float* F, long ldf, long Kp, int Np, int Mp,
                 long K, int N, TO* C, long ldc, int accumulate, int parts, hipStream_t s) {
  if (Kp % BT || Np % BT || Mp % BK || lda % 4 || ldf % 4) return -1;
  BfArgs g{};
  g.P = F; g.ldp = ldf; g.Q = A; g.ldq = lda;
  const bool big = parts == 2 && Kp % BB == 0 && Np % BB == 0;
  const int bt = big ? BB : BT;
  g.tp = Np / bt; g.tq = (int)(Kp / bt);
  g.np = N; g.nq = (int)K;
  g.pfast = 1;
  g.nr = Mp;
  g.rper = Mp;
  g.out = C; g.ldo = ldc; g.sstride = 0; g.accumulate = accumulate;
  g.vec = ((uintptr_t)C % 16 == 0 && ldc % 4 == 0) ? 1 : 0;
  if (big) launchb<2, false, TO>(g, 1, s);
  else if (parts == 3) launch<3, false, TO>(g, 1, s);
  else launch<2, false, TO>(g, 1, s);
  return 0;
}

// the force product Y = A F (A = c, K x Mp, F = W) with the force epilogue:
// fpart[(m*3 + j)*ntq + t] for the ntq = Kp/256 row tiles; 256-tile shapes
// and the 2-part split only (returns -1 otherwise: the caller materializes Y)
int gemm_bf_nn_forces(const float* A, long lda, const float* F, long ldf, long Kp, int Np,
                      int Mp, int N, const float* anl, long lda_anl, const double* gx, int ngw,
                      double* fpart, hipStream_t s) {
  if (Kp % BB || Np % BB || Mp % BK || lda % 4 || ldf % 4 || lda_anl % 4) return -1;
  BfArgs g{};
  g.P = F; g.ldp = ldf; g.Q = A; g.ldq = lda;
  g.tp = Np / BB; g.tq = (int)(Kp / BB);
  g.np = N; g.nq = (int)Kp;
  g.pfast = 1;
  g.nr = Mp;
  g.rper = Mp;
  g.fa = anl; g.lda_fa = lda_anl; g.gx = gx; g.ngw = ngw; g.ntq = g.tq; g.fpart = fpart;
  const unsigned grid = xcd_grid(g, g.tp * g.tq);
  hipLaunchKernelGGL((k_gemm_bfd<false, float, true>), dim3(grid), dim3(512), 0, s, g);
  return 0;
}

int gemm_bf_nn(const float* A, long lda, const float* F, long ldf, long Kp, int Np, int Mp,
               long K, int N, double* C, long ldc, int accumulate, int parts, hipStream_t s) {
  return bf_nn(A, lda, F, ldf, Kp, Np, Mp, K, N, C, ldc, accumulate, parts, s);
}
int gemm_bf_nn(const float* A, long lda, const float* F, long ldf, long Kp, int Np, int Mp,
               long K, int N, float* C, long ldc, int accumulate, int parts, hipStream_t s) {
  return bf_nn(A, lda, F, ldf, Kp, Np, Mp, K, N, C, ldc, accumulate, parts, s);
}

// ------------------------------------------------ FP32 operand staging
// c (FP64 complex, ld ldc complex) -> c32[k + n*ldk], real view, K = 2 ngw;
// columns [nst, nzero) are written as zeros
__global__ void k_to_f32(const double* __restrict__ c, long ldc2, long K, int nst, int nzero,
                         float* __restrict__ c32, long ldk) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= K * nzero) return;
  const long k = t % K, n = t / K;
  c32[k + n * ldk] = n < nst ? (float)c[k + n * ldc2] : 0.0f;
}
void convert_f32(const double* c, long ldc2, long K, int nst, float* c32, long ldk,
                 hipStream_t s, int nzero) {
  nzero = nzero > nst ? nzero : nst;
  const long tot = K * nzero;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, c, ldc2, K,
                     nst, nzero, c32, ldk);
}

// fout (FP64, M x nst, ld M) -> fout32[m + n*ldf] and W32[n + m*ldw] = occ_n fout[m,n];
// band columns [nst, nzero) are written as zeros
__global__ void k_fout_f32(const double* __restrict__ fout, const double* __restrict__ occ, int M,
                           int nst, int nzero, float* __restrict__ f32, long ldf,
                           float* __restrict__ w32, long ldw) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)M * nzero) return;
  const int m = (int)(t % M), n = (int)(t / M);
  const double v = n < nst ? fout[t] : 0.0;
  f32[m + n * ldf] = (float)v;
  if (w32) w32[n + m * ldw] = n < nst ? (float)(occ[n] * v) : 0.0f;
}
void fout_f32(const double* fout, const double* occ, int M, int nst, float* f32, long ldf,
              float* w32, long ldw, hipStream_t s, int nzero) {
  nzero = nzero > nst ? nzero : nst;
  const long tot = (long)M * nzero;
  hipLaunchKernelGGL(k_fout_f32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, fout, occ,
                     M, nst, nzero, f32, ldf, w32, ldw);
}

}  // namespace qbg
