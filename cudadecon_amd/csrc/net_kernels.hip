// Kernels of the warm-start network: a 3-level 3-D U-Net (16/32/64 channels,
// 3x3x3 convolutions, ReLU, average pooling, nearest upsampling with additive
// skips, a 1x1x1 head and a skip add to the normalised input stack).
//
// Activations are BF16, channels-last ([sample][z][y][x][channel]); every
// 3x3x3 convolution is an implicit GEMM on the CDNA4 matrix cores
// (v_mfma 16x16x32 bf16, or 16x16x16 bf16 for 16 input channels) with FP32
// accumulation.  Weights are staged once per workgroup in LDS in the order
// the MFMA A operand is read (one contiguous 8- or 16-byte fragment per lane).
//
// MFMA operand layout (checked on gfx950): for a 16 x 16 x K product with
// F = K/4 values per lane, lane l = 16q + r holds A[m = r][k = F q + i],
// B[k = F q + i][n = r] (i < F) and receives D[m = 4q + i][n = r] (i < 4).
// Convolutions use A = weights (m = output channel, k = input channel),
// B = activations (k = input channel, n = voxel along x), so each lane
// stores 4 consecutive output channels of one voxel.
#include "common.hpp"
#include "net_kernels.hpp"

#include <algorithm>

namespace cdn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even (v_cvt_pk_bf16_f32; inputs are finite)
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16_t f2bf(float f) { return __builtin_bit_cast(bf16_t, (__bf16)f); }
__device__ __forceinline__ unsigned pack2(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t){a, b}, bf16x2_t));
}
// max(x, 0) for finite x as one v_med3_f32 (fmaxf adds a NaN-quieting max)
__device__ __forceinline__ float relu(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 3.4028235e38f); }
__device__ __forceinline__ float lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

template <int KS> struct Mfma;
template <> struct Mfma<32> {
  typedef u32x4 frag;
  __device__ static f32x4 run(frag a, frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
  }
};
template <> struct Mfma<16> {
  typedef u32x2 frag;
  __device__ static f32x4 run(frag a, frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a), __builtin_bit_cast(s16x4, b), c,
                                                     0, 0, 0);
  }
};

__device__ __forceinline__ void tap_offsets(int t, int& dz, int& dy, int& dx) {
  dz = t / 9 - 1;
  dy = (t / 3) % 3 - 1;
  dx = t % 3 - 1;
}

// packed BF16 operand elements of a layer with ci input / co output channels:
// 1 input channel: 32 taps (27 + zeros); 16: 14 tap pairs; else 27 taps
__host__ __device__ constexpr long packed_weights(int ci, int co) {
  return ci == 1 ? 32L * co : (ci == 16 ? 28L * 16 * co : 27L * ci * co);
}

// Store 4 consecutive channels (acc + bias, optional ReLU) of one voxel.
template <bool RELU>
__device__ __forceinline__ u32x2 epi4(f32x4 a, const float* b) {
  float v0 = a[0] + b[0], v1 = a[1] + b[1], v2 = a[2] + b[2], v3 = a[3] + b[3];
  if (RELU) {
    v0 = relu(v0);
    v1 = relu(v1);
    v2 = relu(v2);
    v3 = relu(v3);
  }
  u32x2 w;
  w[0] = pack2(v0, v1);
  w[1] = pack2(v2, v3);
  return w;
}
template <bool RELU>
__device__ __forceinline__ void store4(bf16_t* p, f32x4 a, const float* b) {
  *reinterpret_cast<u32x2*>(p) = epi4<RELU>(a, b);
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, CIN in {16, 32, 64}, zero padding, per-sample borders.
// One wave computes 16*MB consecutive voxels of one (sample, z, y) row for
// all COUT channels.  Grid-stride over (row, x-chunk) work items.
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int MB, int THREADS, bool RELU, bool BIAS>
__global__ __launch_bounds__(THREADS) void k_conv(const bf16_t* __restrict__ in, const bf16_t* __restrict__ wf,
                                                  const float* __restrict__ bias, bf16_t* __restrict__ out, Vol v) {
  constexpr int KS = CIN == 16 ? 16 : 32, NS = CIN / KS, NBO = COUT / 16, FR = KS / 4;
  constexpr int WN = 27 * CIN * COUT, WAVES = THREADS / 64;
  typedef typename Mfma<KS>::frag frag;
  __shared__ __attribute__((aligned(16))) bf16_t sw[WN];
  for (int i = threadIdx.x * 8; i < WN; i += THREADS * 8)
    *reinterpret_cast<u32x4*>(sw + i) = *reinterpret_cast<const u32x4*>(wf + i);
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  float bl[NBO][4];
#pragma unroll
  for (int j = 0; j < NBO; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) bl[j][i] = BIAS ? bias[16 * j + 4 * q + i] : 0.f;
  const int xchunks = (v.nx + 16 * MB - 1) / (16 * MB);
  const long rows = (long)v.nb * v.nz * v.ny;
  const long items = rows * xchunks;
  for (long it = (long)blockIdx.x * WAVES + (threadIdx.x >> 6); it < items; it += (long)gridDim.x * WAVES) {
    const long row = it / xchunks;
    const int x0 = (int)(it - row * xchunks) * 16 * MB;
    const int y = (int)(row % v.ny);
    const long bzi = row / v.ny;
    const int z = (int)(bzi % v.nz);
    const long sbase = (bzi - z) * v.ny;
    f32x4 acc[NBO][MB];
#pragma unroll
    for (int j = 0; j < NBO; ++j)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) acc[j][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
      const int zz = z + dz;
      if (zz < 0 || zz >= v.nz) continue;
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= v.ny) continue;
        const bf16_t* src = in + (sbase + (long)zz * v.ny + yy) * (long)v.nx * CIN;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
#pragma unroll
          for (int s = 0; s < NS; ++s) {
            frag b[MB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
              const int x = x0 + 16 * mb + r + dx;
              if (x >= 0 && x < v.nx)
                b[mb] = *reinterpret_cast<const frag*>(src + (long)x * CIN + s * KS + q * FR);
              else
                b[mb] = frag(0u);
            }
#pragma unroll
            for (int j = 0; j < NBO; ++j) {
              const frag a = *reinterpret_cast<const frag*>(sw + (((t * NS + s) * NBO + j) * 64 + lane) * FR);
#pragma unroll
              for (int mb = 0; mb < MB; ++mb) acc[j][mb] = Mfma<KS>::run(a, b[mb], acc[j][mb]);
            }
          }
        }
      }
    }
    bf16_t* dst = out + row * (long)v.nx * COUT;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int x = x0 + 16 * mb + r;
      if (x < v.nx)
#pragma unroll
        for (int j = 0; j < NBO; ++j) store4<RELU>(dst + (long)x * COUT + 16 * j + 4 * q, acc[j][mb], bl[j]);
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution with 2.5-D blocking.  A workgroup of 4 waves owns 4
// output rows (wave w: row y0 + w) of TX voxels and walks a chunk of
// z-planes.  The input planes z-1, z, z+1 (rows y0-1..y0+4, columns
// x0-1..x0+TX, all CIN channels, zero outside the volume) sit in an LDS ring
// of 3 planes, so every tap reads its B fragment from LDS without branches;
// plane z+2 is loaded into registers while plane z is computed and stored
// into the slot of plane z-1 after it.  Weights (A fragments) are read from
// the packed array (L1/L2 resident).  Same arithmetic as k_conv.
// ---------------------------------------------------------------------------
// UPADD: the input is up(coarse) + in (nearest upsampling of a half-size
// tensor of the same channels), formed while the planes are loaded, as
// k_upadd stores it (bf16 of the FP32 sum).
// POOL: also writes pooled = 2x2x2 mean of the output, as k_pool computes it
// from the stored values (same children order, FP32 sums).  The bf16 outputs
// of a z pair go to an LDS buffer; after the pair's second plane each thread
// forms one 8-channel group of one coarse voxel of the tile (z chunks start
// on even planes, so pairs never straddle chunks).
template <int CIN, int COUT, int TX, bool RELU, bool BIAS, bool UPADD, bool POOL = false>
__global__ __launch_bounds__(256) void k_conv25(const bf16_t* __restrict__ in, const bf16_t* __restrict__ wf,
                                                const float* __restrict__ bias, bf16_t* __restrict__ out, Vol v,
                                                int zchunk, const bf16_t* __restrict__ coarse,
                                                bf16_t* __restrict__ pooled) {
  constexpr int KS = CIN == 16 ? 16 : 32, NS = CIN / KS, NBO = COUT / 16, FR = KS / 4;
  constexpr int TY = 4, MB = TX / 16, RW = TX + 2, ROWS = TY + 2;
  // one output-channel block and one K step: the waves split the tile by
  // column blocks instead of rows
  constexpr bool COLBLK = CIN == 32 && COUT == 16 && !POOL;
  static_assert(!COLBLK || MB == TY, "one column block per wave");
  constexpr int ROWCH = RW * CIN / 8;          // 16-byte chunks per tile row
  constexpr int PLANE = ROWS * RW * CIN;       // bf16 per plane
  constexpr int CHUNKS = ROWS * ROWCH;         // 16-byte chunks per plane
  constexpr int PER_T = (CHUNKS + 255) / 256;  // chunks per thread
  typedef typename Mfma<KS>::frag frag;
  __shared__ __attribute__((aligned(16))) bf16_t ring[3 * PLANE];
  __shared__ __attribute__((aligned(16))) bf16_t pool_buf[POOL ? 2 * TY * TX * COUT : 8];
  static_assert(!POOL || TX * COUT / 8 == 256, "POOL: one 8-channel coarse group per thread");
  const int xt_n = (v.nx + TX - 1) / TX, yt_n = (v.ny + TY - 1) / TY, zc_n = (v.nz + zchunk - 1) / zchunk;
  int blk = blockIdx.x;
  const int xt = blk % xt_n;
  blk /= xt_n;
  const int yt = blk % yt_n;
  blk /= yt_n;
  const int zc = blk % zc_n;
  const int b = blk / zc_n;
  const int x0 = xt * TX, y0 = yt * TY, z0 = zc * zchunk, z1 = min(z0 + zchunk, v.nz);
  const bf16_t* smp = in + (long)b * v.nz * v.ny * v.nx * CIN;
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, w = threadIdx.x >> 6;

  u32x4 reg[PER_T];
  auto load_plane = [&](int zz) {
#pragma unroll
    for (int k = 0; k < PER_T; ++k) {
      const int c = threadIdx.x + 256 * k;
      const int rr = c / ROWCH, rem = c - rr * ROWCH;
      const int xx = rem / (CIN / 8), cg = rem - xx * (CIN / 8);
      const int y = y0 - 1 + rr, x = x0 - 1 + xx;
      const bool ok = c < CHUNKS && zz >= 0 && zz < v.nz && y >= 0 && y < v.ny && x >= 0 && x < v.nx;
      const long off = ok ? ((((long)zz * v.ny + y) * v.nx + x) * CIN + cg * 8) : 0L;
      u32x4 d = *reinterpret_cast<const u32x4*>(smp + off);
      if constexpr (UPADD) {
        const int hz = v.nz / 2, hy = v.ny / 2, hx = v.nx / 2;
        const long coff = ok ? (((((long)b * hz + zz / 2) * hy + y / 2) * hx + x / 2) * CIN + cg * 8) : 0L;
        const u32x4 c = *reinterpret_cast<const u32x4*>(coarse + coff);
#pragma unroll
        for (int e = 0; e < 4; ++e) d[e] = pack2(lo(c[e]) + lo(d[e]), hi(c[e]) + hi(d[e]));
      }
      reg[k] = ok ? d : u32x4(0u);
    }
  };
  auto store_plane = [&](int zz) {
    bf16_t* dst = ring + ((zz + 3) % 3) * PLANE;
#pragma unroll
    for (int k = 0; k < PER_T; ++k) {
      const int c = threadIdx.x + 256 * k;
      if (c < CHUNKS) *reinterpret_cast<u32x4*>(dst + c * 8) = reg[k];
    }
  };
  float bl[NBO][4];
#pragma unroll
  for (int j = 0; j < NBO; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) bl[j][i] = BIAS ? bias[16 * j + 4 * q + i] : 0.f;

  load_plane(z0 - 1);
  store_plane(z0 - 1);
  load_plane(z0);
  store_plane(z0);
  load_plane(z0 + 1);
  store_plane(z0 + 1);
  const int y = y0 + w;
  for (int z = z0; z < z1; ++z) {
    if (z + 2 <= z1) load_plane(z + 2);  // in flight while plane z is computed
    __syncthreads();
    if constexpr (COLBLK) {
      // 32 -> 16: wave w owns columns x0 + 16w .. + 15 of all TY rows, so the
      // B fragment of an input row (read once from LDS) feeds the up to 3
      // output rows it is a tap of; taps accumulate in (dz, dx, dy) order
      f32x4 acc4[TY];
#pragma unroll
      for (int j = 0; j < TY; ++j) acc4[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        const bf16_t* base = ring + ((z + dz + 3) % 3) * PLANE;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          frag a[3];
#pragma unroll
          for (int dy = -1; dy <= 1; ++dy) {
            const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
            a[dy + 1] = *reinterpret_cast<const frag*>(wf + (t * 64 + lane) * FR);
          }
#pragma unroll
          for (int rr = 0; rr < ROWS; ++rr) {
            const frag bfr = *reinterpret_cast<const frag*>(base + (rr * RW + 16 * w + r + 1 + dx) * CIN + q * FR);
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
              const int j = rr - 1 - dy;
              if (j >= 0 && j < TY) acc4[j] = Mfma<KS>::run(a[dy + 1], bfr, acc4[j]);
            }
          }
        }
      }
      const int x = x0 + 16 * w + r;
#pragma unroll
      for (int j = 0; j < TY; ++j) {
        const int yj = y0 + j;
        if (yj < v.ny && x < v.nx)
          store4<RELU>(out + ((((long)b * v.nz + z) * v.ny + yj) * v.nx + x) * COUT + 4 * q, acc4[j], bl[0]);
      }
    } else {
      f32x4 acc[NBO][MB];
  #pragma unroll
      for (int j = 0; j < NBO; ++j)
  #pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[j][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
      if constexpr (CIN == 16) {
        // 14 tap pairs on the full-rate K = 32 instruction: lanes q < 2 read
        // tap 2p, lanes q >= 2 tap 2p + 1 (tap 27 has zero weights)
        int slot[3];
  #pragma unroll
        for (int d = 0; d < 3; ++d) slot[d] = ((z + d - 1 + 3) % 3) * PLANE;
        const bool hi = q >= 2;
        const int cofs = 8 * (q & 1);
  #pragma unroll
        for (int pr = 0; pr < 14; ++pr) {
          const int t0 = 2 * pr, t1 = 2 * pr + 1 < 27 ? 2 * pr + 1 : 26;
          const int o0 = slot[t0 / 9] + ((w + (t0 / 3) % 3) * RW + t0 % 3) * 16;
          const int o1 = slot[t1 / 9] + ((w + (t1 / 3) % 3) * RW + t1 % 3) * 16;
          const bf16_t* base = ring + (hi ? o1 : o0) + cofs;
          u32x4 bfr[MB];
  #pragma unroll
          for (int mb = 0; mb < MB; ++mb) bfr[mb] = *reinterpret_cast<const u32x4*>(base + (16 * mb + r) * 16);
  #pragma unroll
          for (int j = 0; j < NBO; ++j) {
            const u32x4 a = *reinterpret_cast<const u32x4*>(wf + ((pr * NBO + j) * 64 + lane) * 8);
  #pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc[j][mb] = Mfma<32>::run(a, bfr[mb], acc[j][mb]);
          }
        }
      } else {
      // 64-channel layers and the fused up-add: the dz loop is not unrolled (an
      // unrolled one keeps too many fragments live: spills or occupancy 1)
      constexpr int kDzUnroll = (CIN == 64 || COUT == 64 || UPADD) ? 1 : 3;
  #pragma unroll kDzUnroll
      for (int dz = -1; dz <= 1; ++dz) {
        const bf16_t* base = ring + ((z + dz + 3) % 3) * PLANE;
  #pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
  #pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
  #pragma unroll
            for (int s = 0; s < NS; ++s) {
              frag bfr[MB];
  #pragma unroll
              for (int mb = 0; mb < MB; ++mb)
                bfr[mb] = *reinterpret_cast<const frag*>(base + ((w + 1 + dy) * RW + 16 * mb + r + 1 + dx) * CIN +
                                                         s * KS + q * FR);
  #pragma unroll
              for (int j = 0; j < NBO; ++j) {
                const frag a = *reinterpret_cast<const frag*>(wf + (((t * NS + s) * NBO + j) * 64 + lane) * FR);
  #pragma unroll
                for (int mb = 0; mb < MB; ++mb) acc[j][mb] = Mfma<KS>::run(a, bfr[mb], acc[j][mb]);
              }
            }
          }
        }
      }
      }
      if constexpr (POOL) {
        // every row and column of the tile goes to the pool buffer (the ones
        // outside the volume hold finite values that no output uses)
        bf16_t* pb = pool_buf + (((z - z0) & 1) * TY + w) * TX * COUT;
        bf16_t* dst = out + ((((long)b * v.nz + z) * v.ny + (y < v.ny ? y : 0)) * v.nx) * COUT;
  #pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const int x = x0 + 16 * mb + r;
  #pragma unroll
          for (int j = 0; j < NBO; ++j) {
            const u32x2 e = epi4<RELU>(acc[j][mb], bl[j]);
            *reinterpret_cast<u32x2*>(pb + (16 * mb + r) * COUT + 16 * j + 4 * q) = e;
            if (y < v.ny && x < v.nx) *reinterpret_cast<u32x2*>(dst + (long)x * COUT + 16 * j + 4 * q) = e;
          }
        }
      } else if (y < v.ny) {
        bf16_t* dst = out + ((((long)b * v.nz + z) * v.ny + y) * v.nx) * COUT;
  #pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const int x = x0 + 16 * mb + r;
          if (x < v.nx)
  #pragma unroll
            for (int j = 0; j < NBO; ++j) store4<RELU>(dst + (long)x * COUT + 16 * j + 4 * q, acc[j][mb], bl[j]);
        }
      }
    }
    __syncthreads();                     // every wave is done with plane z-1
    if constexpr (POOL) {
      if ((z - z0) & 1) {
        constexpr int G = COUT / 8;
        const int cg = threadIdx.x % G, cxl = (threadIdx.x / G) % (TX / 2), cyl = threadIdx.x / G / (TX / 2);
        float sm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int row = (c >> 2) * TY + 2 * cyl + ((c >> 1) & 1), col = 2 * cxl + (c & 1);
          const u32x4 d = *reinterpret_cast<const u32x4*>(pool_buf + (row * TX + col) * COUT + 8 * cg);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            sm[2 * k] += lo(d[k]);
            sm[2 * k + 1] += hi(d[k]);
          }
        }
        const int hy = v.ny / 2, hx = v.nx / 2, cy = (y0 >> 1) + cyl, cx = (x0 >> 1) + cxl;
        if (cy < hy && cx < hx) {
          u32x4 o;
#pragma unroll
          for (int k = 0; k < 4; ++k) o[k] = pack2(sm[2 * k] * 0.125f, sm[2 * k + 1] * 0.125f);
          const long ow = (((long)b * (v.nz / 2) + (z >> 1)) * hy + cy) * (long)hx + cx;
          *reinterpret_cast<u32x4*>(pooled + ow * COUT + 8 * cg) = o;
        }
      }
    }
    if (z + 2 <= z1) store_plane(z + 2);  // into the slot of plane z-1
  }
}

// First layer: 1 -> 16 channels from the FP32 stack scaled by 1/s per sample,
// K running over the 27 taps (padded to 32 with zero weights).  A workgroup
// of 4 waves owns rows y0..y0+7 and columns x0..x0+63 (wave w: x0 + 16w ..
// x0 + 16w + 15, one M-block per row) and walks the planes zb..ze-1 of its
// z-chunk.  LDS holds bf16(raw / s) of 4 planes over rows y0-1..y0+8 and
// columns x0-1..x0+64 (0 outside the volume): plane p in slot
// (p - zb + 1) mod 4, slots 0 and 1 mirrored in slots 4 and 5, so planes
// z-1, z, z+1 are always the consecutive slots from (z - zb) mod 4 and each
// lane reads its 8 taps at fixed offsets from one base per plane.  While
// plane z is computed, plane z+2 is stored and plane z+3 is in flight to
// registers; one barrier per plane orders the ring.
// MODE 0: write e1;  1: write pool(e1) only;  2: fused head, write y;
// 3: fused head, write the RL seed max(y, 0) * s.  In the head modes e1 is
// recomputed here and never stored; d2 and the head parameters hp[0..16]
// complete y = u + bh + sum_c Wh[c] (e1[c] + d2[v/2][c]).
constexpr int FT_X = 66, FT_Y = 10, FT_PLANE = FT_Y * FT_X, FT_SLOTS = 6;

template <int MODE>
__global__ __launch_bounds__(256) void k_first(const float* __restrict__ raw, const float* __restrict__ inv_s,
                                               const bf16_t* __restrict__ wf, const float* __restrict__ bias,
                                               bf16_t* __restrict__ out16, const bf16_t* __restrict__ d2,
                                               const float* __restrict__ hp, float* __restrict__ out32, Vol v,
                                               int zchunk) {
  __shared__ bf16_t tile[FT_SLOTS * FT_PLANE];
  const int xb = (v.nx + 63) / 64, yb = (v.ny + 7) / 8, zn = (v.nz + zchunk - 1) / zchunk;
  int blk = blockIdx.x;
  const int xt = blk % xb;
  blk /= xb;
  const int yt = blk % yb;
  blk /= yb;
  const int zt = blk % zn;
  const int b = blk / zn;
  const int y0 = 8 * yt, x0 = 64 * xt, zb = zt * zchunk;
  const int ze = min(zb + zchunk, v.nz);
  const float sc = inv_s[b], smean = 1.f / sc;
  const long plane = (long)v.ny * v.nx;
  const float* smp = raw + (long)b * v.nz * plane;
  // loader: tile element e = tid + 256k (k < 3, e < 660) of a plane, row
  // e / 66, column e % 66; in-plane offset and validity fixed per thread
  int loff[3];
  bool lok[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int e = threadIdx.x + 256 * k, ry = e / FT_X, cx = e - ry * FT_X;
    const int y = y0 - 1 + ry, x = x0 - 1 + cx;
    lok[k] = e < FT_PLANE && y >= 0 && y < v.ny && x >= 0 && x < v.nx;
    loff[k] = lok[k] ? y * v.nx + x : 0;
  }
  auto load_plane = [&](int z, float (&val)[3]) {
    const bool zok = z >= 0 && z < v.nz;
    const float* pl = smp + (zok ? z : 0) * plane;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float f = pl[loff[k]];
      val[k] = zok && lok[k] ? f * sc : 0.f;
    }
  };
  auto store_plane = [&](int z, const float (&val)[3]) {
    const int slot = (z - zb + 1) & 3;
    bf16_t* sl = tile + slot * FT_PLANE;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e < FT_PLANE) {
        const bf16_t h = f2bf(val[k]);
        sl[e] = h;
        if (slot < 2) sl[4 * FT_PLANE + e] = h;
      }
    }
  };
  {
    float p0[3], p1[3], p2[3];
    load_plane(zb - 1, p0);
    load_plane(zb, p1);
    load_plane(zb + 1, p2);
    store_plane(zb - 1, p0);
    store_plane(zb, p1);
    store_plane(zb + 1, p2);
  }
  float nxt[3] = {0.f, 0.f, 0.f};
  if (zb + 2 <= ze) load_plane(zb + 2, nxt);
  __syncthreads();

  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15, w = threadIdx.x >> 6;
  const u32x4 a = *reinterpret_cast<const u32x4*>(wf + lane * 8);
  // tap offsets within the 3 consecutive slots of planes z-1..z+1; the
  // padding taps 27..31 read tap 26's element (their weights are 0)
  int toff[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int t = 8 * q + i;
    int dz, dy, dx;
    tap_offsets(t < 27 ? t : 26, dz, dy, dx);
    toff[i] = (1 + dz) * FT_PLANE + (1 + dy) * FT_X + (1 + dx) + 16 * w + r;
  }
  f32x4 bl;
#pragma unroll
  for (int i = 0; i < 4; ++i) bl[i] = bias[4 * q + i];
  const int x = x0 + 16 * w + r;
  const bool xin = x < v.nx;
  float wh[4] = {0.f, 0.f, 0.f, 0.f}, bh = 0.f;
  if (MODE >= 2) {
#pragma unroll
    for (int i = 0; i < 4; ++i) wh[i] = hp[4 * q + i];
    bh = hp[16];
  }
  const int hz = v.nz / 2;
  float ps[4][4];  // MODE 1: pooled sums per y pair over a z pair
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int i = 0; i < 4; ++i) ps[k][i] = 0.f;
  float hd[4];  // head modes: sum_c Wh[c] d2[c] of the current z pair's coarse voxels, per y pair

  // head modes: after the 8 rows, lane (q, r) owns rows hrow(q, 0..1) of
  // column x (see the transposed reduction below)
  const int hrow0 = 4 * (q & 1) + 2 * (q >> 1);
  for (int z = zb; z < ze; ++z) {
    const bf16_t* tz = tile + ((z - zb) & 3) * FT_PLANE;
    float uin[2];
    if (MODE >= 2) {
      if (((z - zb) & 1) == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int yc = (y0 >> 1) + k;
          const bool ok = xin && 2 * yc < v.ny;
          const long cw = (((long)b * hz + (z >> 1)) * (v.ny / 2) + (ok ? yc : 0)) * (v.nx / 2) + (ok ? (x >> 1) : 0);
          u32x2 d = *reinterpret_cast<const u32x2*>(d2 + cw * 16 + 4 * q);
          d = ok ? d : u32x2(0u);
          hd[k] = wh[0] * lo(d[0]);
          hd[k] += wh[1] * hi(d[0]);
          hd[k] += wh[2] * lo(d[1]);
          hd[k] += wh[3] * hi(d[1]);
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int y = y0 + hrow0 + j;
        const bool ok = xin && y < v.ny;
        const float f = smp[z * plane + (ok ? y : 0) * v.nx + (ok ? x : 0)];
        uin[j] = ok ? f * sc : 0.f;
      }
    }
    auto gather = [&](int my) {
      const bf16_t* tb = tz + my * FT_X;
      u32x4 bf;
#pragma unroll
      for (int i = 0; i < 8; i += 2) bf[i / 2] = (unsigned)tb[toff[i]] | ((unsigned)tb[toff[i + 1]] << 16);
      return bf;
    };
    u32x2 o16[8];    // MODE 0: e1 of the 8 rows
    float part[8];   // head modes: this lane's 4-channel share of sum_c Wh[c] (e1 + d2)
    u32x4 bcur = gather(0);
#pragma unroll
    for (int my = 0; my < 8; ++my) {
      const f32x4 acc = Mfma<32>::run(a, bcur, bl);  // bias as the accumulator input
      if (my < 7) bcur = gather(my + 1);              // next row's operand while the MFMA runs
      const unsigned e01 = pack2(relu(acc[0]), relu(acc[1]));
      const unsigned e23 = pack2(relu(acc[2]), relu(acc[3]));
      if (MODE == 0) {
        o16[my] = u32x2{e01, e23};
        continue;
      }
      const float e[4] = {lo(e01), hi(e01), lo(e23), hi(e23)};  // e1 as stored by MODE 0
      if (MODE == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) ps[my >> 1][i] += e[i];
        continue;
      }
      // this lane's 4 channels of sum_c Wh[c] (e1[c] + d2[c])
      float pt = wh[0] * e[0];
#pragma unroll
      for (int i = 1; i < 4; ++i) pt += wh[i] * e[i];
      part[my] = pt + hd[my >> 1];
    }
    if (MODE == 0) {
#pragma unroll
      for (int my = 0; my < 8; ++my) {
        const int y = y0 + my;
        if (xin && y < v.ny) {
          const long vox = (((long)b * v.nz + z) * v.ny + y) * v.nx + x;
          *reinterpret_cast<u32x2*>(out16 + vox * 16 + 4 * q) = o16[my];
        }
      }
    }
    if (MODE >= 2) {
      // sum over the 4 lane groups, transposed: xor 16 pairs groups (0,1) and
      // (2,3), even groups keeping rows 0..3 and odd ones rows 4..7; xor 32
      // then leaves group q rows hrow0, hrow0 + 1.  Each total is
      // (g0 + g1) + (g2 + g3) in this order.
      float t4[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool odd = q & 1;
        const float keep = odd ? part[4 + j] : part[j], give = odd ? part[j] : part[4 + j];
        t4[j] = keep + __shfl_xor(give, 16);
      }
      float tot[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const bool hi2 = q >> 1;
        const float keep = hi2 ? t4[2 + j] : t4[j], give = hi2 ? t4[j] : t4[2 + j];
        tot[j] = keep + __shfl_xor(give, 32);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int y = y0 + hrow0 + j;
        if (xin && y < v.ny) {
          const long vox = (((long)b * v.nz + z) * v.ny + y) * v.nx + x;
          const float yv = uin[j] + bh + tot[j];
          out32[vox] = MODE == 2 ? yv : relu(yv) * smean;
        }
      }
    }
    if (MODE == 1 && ((z - zb) & 1) == 1) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) ps[k][i] += __shfl_xor(ps[k][i], 1);
        const int yc = (y0 >> 1) + k;
        if (xin && (r & 1) == 0 && 2 * yc < v.ny) {
          const long cw = (((long)b * hz + (z >> 1)) * (v.ny / 2) + yc) * (v.nx / 2) + (x >> 1);
          u32x2 o;
          o[0] = pack2(0.125f * ps[k][0], 0.125f * ps[k][1]);
          o[1] = pack2(0.125f * ps[k][2], 0.125f * ps[k][3]);
          *reinterpret_cast<u32x2*>(out16 + cw * 16 + 4 * q) = o;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) ps[k][i] = 0.f;
      }
    }
    // plane z+2 into the slot of z-2 (last read before the previous barrier),
    // plane z+3 into registers
    if (z + 2 <= ze) store_plane(z + 2, nxt);
    if (z + 3 <= ze) load_plane(z + 3, nxt);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pooling / upsampling (level L has dims v; the coarser level has v/2)
// ---------------------------------------------------------------------------
// p[w] = mean of the 8 children of w; one thread per (coarse voxel, 8 channels)
template <int C>
__global__ void k_pool(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, Vol fine) {
  constexpr int G = C / 8;
  const int cz = fine.nz / 2, cy = fine.ny / 2, cx = fine.nx / 2;
  const long n = (long)fine.nb * cz * cy * cx * G;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int g = (int)(i % G);
    long w = i / G;
    const int x = (int)(w % cx);
    w /= cx;
    const int y = (int)(w % cy);
    w /= cy;
    const int z = (int)(w % cz);
    const long b = w / cz;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const long f = ((b * fine.nz + 2 * z + (c >> 2)) * fine.ny + 2 * y + ((c >> 1) & 1)) * fine.nx + 2 * x + (c & 1);
      const u32x4 d = *reinterpret_cast<const u32x4*>(in + f * C + 8 * g);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += lo(d[k]);
        s[2 * k + 1] += hi(d[k]);
      }
    }
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = pack2(s[2 * k] * 0.125f, s[2 * k + 1] * 0.125f);
    const long ow = ((b * cz + z) * cy + y) * (long)cx + x;
    *reinterpret_cast<u32x4*>(out + ow * C + 8 * g) = o;
  }
}

// out[v] = coarse[v/2] + skip[v]
template <int C>
__global__ void k_upadd(const bf16_t* __restrict__ coarse, const bf16_t* __restrict__ skip, bf16_t* __restrict__ out,
                        Vol fine) {
  constexpr int G = C / 8;
  const long n = (long)fine.nb * fine.nz * fine.ny * fine.nx * G;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int g = (int)(i % G);
    long w = i / G;
    const long vox = w;
    const int x = (int)(w % fine.nx);
    w /= fine.nx;
    const int y = (int)(w % fine.ny);
    w /= fine.ny;
    const int z = (int)(w % fine.nz);
    const long b = w / fine.nz;
    const long cw = ((b * (fine.nz / 2) + z / 2) * (fine.ny / 2) + y / 2) * (long)(fine.nx / 2) + x / 2;
    const u32x4 a = *reinterpret_cast<const u32x4*>(coarse + cw * C + 8 * g);
    const u32x4 s = *reinterpret_cast<const u32x4*>(skip + vox * C + 8 * g);
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = pack2(lo(a[k]) + lo(s[k]), hi(a[k]) + hi(s[k]));
    *reinterpret_cast<u32x4*>(out + vox * C + 8 * g) = o;
  }
}

// ---------------------------------------------------------------------------
// Head: y = u + bh + sum_c Wh[c] (d2[v/2][c] + e1[v][c]),  u = raw / s.
// MODE 0: write y;  MODE 1: write max(y, 0) * s (the RL seed);
// MODE 2 (training): residual against target/s, dL/dy = 2 res / n_total,
// per-block partials of [sum res^2, dWh[16], dbh].
// ---------------------------------------------------------------------------
constexpr int HEAD_THREADS = 256;
constexpr int HEAD_PART = 18;

template <int MODE>
__global__ __launch_bounds__(HEAD_THREADS) void k_head(const float* __restrict__ raw,
                                                       const float* __restrict__ inv_s,
                                                       const bf16_t* __restrict__ e1, const bf16_t* __restrict__ d2,
                                                       const float* __restrict__ hp, const float* __restrict__ tgt,
                                                       float* __restrict__ out, float* __restrict__ part, Vol v,
                                                       float inv_total) {
  float wh[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) wh[c] = hp[c];
  const float bh = hp[16];
  float acc[HEAD_PART];
#pragma unroll
  for (int k = 0; k < HEAD_PART; ++k) acc[k] = 0.f;
  const long per = (long)v.nz * v.ny * v.nx;
  const long n = per * v.nb;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    long w = i;
    const int x = (int)(w % v.nx);
    w /= v.nx;
    const int y = (int)(w % v.ny);
    w /= v.ny;
    const int z = (int)(w % v.nz);
    const long b = w / v.nz;
    const long cw = ((b * (v.nz / 2) + z / 2) * (v.ny / 2) + y / 2) * (long)(v.nx / 2) + x / 2;
    const u32x4* pe = reinterpret_cast<const u32x4*>(e1 + i * 16);
    const u32x4* pd = reinterpret_cast<const u32x4*>(d2 + cw * 16);
    float a1[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const u32x4 ee = pe[h], dd = pd[h];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a1[8 * h + 2 * k] = lo(ee[k]) + lo(dd[k]);
        a1[8 * h + 2 * k + 1] = hi(ee[k]) + hi(dd[k]);
      }
    }
    const float sc = inv_s[b];
    float yv = raw[i] * sc + bh;
#pragma unroll
    for (int c = 0; c < 16; ++c) yv += wh[c] * a1[c];
    if (MODE == 0) out[i] = yv;
    if (MODE == 1) out[i] = fmaxf(yv, 0.f) / sc;
    if (MODE == 2) {
      const float res = yv - tgt[i] * sc;
      const float g = 2.f * res * inv_total;
      out[i] = g;
      acc[0] += res * res;
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[1 + c] += g * a1[c];
      acc[17] += g;
    }
  }
  if (MODE == 2) {
    __shared__ float red[HEAD_THREADS];
    for (int k = 0; k < HEAD_PART; ++k) {
      red[threadIdx.x] = acc[k];
      __syncthreads();
      for (int s = HEAD_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) part[blockIdx.x * HEAD_PART + k] = red[0];
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Backward elementwise kernels (one thread per voxel and 8 channels)
// ---------------------------------------------------------------------------
// Encoder level: dPre = (skip + up(dP)/8) * [e > 0]; skip is a tensor, or the
// rank-1 head gradient dout[v] * Wh[c] when skip == nullptr.  dP may be null.
template <int C>
__global__ void k_grad_enc(const bf16_t* __restrict__ e, const bf16_t* __restrict__ skip,
                           const float* __restrict__ dout, const float* __restrict__ wh,
                           const bf16_t* __restrict__ dP, bf16_t* __restrict__ dpre, Vol v) {
  constexpr int G = C / 8;
  const long n = (long)v.nb * v.nz * v.ny * v.nx * G;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int g = (int)(i % G);
    const long vox = i / G;
    float s[8];
    if (skip) {
      const u32x4 d = *reinterpret_cast<const u32x4*>(skip + vox * C + 8 * g);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] = lo(d[k]);
        s[2 * k + 1] = hi(d[k]);
      }
    } else {
      const float go = dout[vox];
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] = go * wh[8 * g + k];
    }
    if (dP) {
      long w = vox;
      const int x = (int)(w % v.nx);
      w /= v.nx;
      const int y = (int)(w % v.ny);
      w /= v.ny;
      const int z = (int)(w % v.nz);
      const long b = w / v.nz;
      const long cw = ((b * (v.nz / 2) + z / 2) * (v.ny / 2) + y / 2) * (long)(v.nx / 2) + x / 2;
      const u32x4 d = *reinterpret_cast<const u32x4*>(dP + cw * C + 8 * g);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += 0.125f * lo(d[k]);
        s[2 * k + 1] += 0.125f * hi(d[k]);
      }
    }
    const u32x4 m = *reinterpret_cast<const u32x4*>(e + vox * C + 8 * g);
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = pack2(lo(m[k]) > 0.f ? s[2 * k] : 0.f, hi(m[k]) > 0.f ? s[2 * k + 1] : 0.f);
    *reinterpret_cast<u32x4*>(dpre + vox * C + 8 * g) = o;
  }
}

// Decoder level (coarse dims v): dPre[w] = (sum over the 8 children of the
// finer gradient) * [d > 0].  The finer gradient is a tensor, or the rank-1
// head gradient dout * Wh when fine == nullptr.
template <int C>
__global__ void k_grad_dec(const bf16_t* __restrict__ d, const bf16_t* __restrict__ fine,
                           const float* __restrict__ dout, const float* __restrict__ wh, bf16_t* __restrict__ dpre,
                           Vol v) {
  constexpr int G = C / 8;
  const long n = (long)v.nb * v.nz * v.ny * v.nx * G;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int g = (int)(i % G);
    const long vox = i / G;
    long w = vox;
    const int x = (int)(w % v.nx);
    w /= v.nx;
    const int y = (int)(w % v.ny);
    w /= v.ny;
    const int z = (int)(w % v.nz);
    const long b = w / v.nz;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float so = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const long f = ((b * 2 * v.nz + 2 * z + (c >> 2)) * (2 * v.ny) + 2 * y + ((c >> 1) & 1)) * (long)(2 * v.nx) +
                     2 * x + (c & 1);
      if (fine) {
        const u32x4 q = *reinterpret_cast<const u32x4*>(fine + f * C + 8 * g);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          s[2 * k] += lo(q[k]);
          s[2 * k + 1] += hi(q[k]);
        }
      } else {
        so += dout[f];
      }
    }
    if (!fine)
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] = so * wh[8 * g + k];
    const u32x4 m = *reinterpret_cast<const u32x4*>(d + vox * C + 8 * g);
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = pack2(lo(m[k]) > 0.f ? s[2 * k] : 0.f, hi(m[k]) > 0.f ? s[2 * k + 1] : 0.f);
    *reinterpret_cast<u32x4*>(dpre + vox * C + 8 * g) = o;
  }
}

// Per-block channel sums of a [n][C] BF16 tensor (bias gradients):
// part[block][C].  Threads own a fixed channel group; LDS tree per group.
template <int C>
__global__ __launch_bounds__(256) void k_chan_sum(const bf16_t* __restrict__ t, long nvox, float* __restrict__ part) {
  constexpr int G = C / 8;
  const int g = threadIdx.x % G;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const long n = nvox * G;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const u32x4 d = *reinterpret_cast<const u32x4*>(t + (i / G) * C + 8 * g);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      s[2 * k] += lo(d[k]);
      s[2 * k + 1] += hi(d[k]);
    }
  }
  __shared__ float red[256][9];
#pragma unroll
  for (int k = 0; k < 8; ++k) red[threadIdx.x][k] = s[k];
  __syncthreads();
  // 256 is a multiple of G: threads tid, tid+G, ... share a channel group
  for (int stride = 128; stride >= G; stride >>= 1) {
    if ((int)threadIdx.x < stride)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[threadIdx.x][k] += red[threadIdx.x + stride][k];
    __syncthreads();
  }
  if ((int)threadIdx.x < G)
#pragma unroll
    for (int k = 0; k < 8; ++k) part[blockIdx.x * (long)C + 8 * threadIdx.x + k] = red[threadIdx.x][k];
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[t][ci][co] = sum_v in[v + t][ci] * dPre[v][co].
// MFMA with K = 32 voxels along x: A = dPre^T (m = co), B = shifted input
// (n = ci).  Wave (tap, chunk) accumulates over work items chunk, chunk+G, ...
// and writes part[chunk][t][ci][co].
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(const bf16_t* __restrict__ in, const bf16_t* __restrict__ gr, Vol v,
                                               int G, float* __restrict__ part) {
  constexpr int NBO = COUT / 16, NBI = CIN / 16;
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wv >= 27 * G) return;
  const int t = wv % 27, chunk = wv / 27;
  int dz, dy, dx;
  tap_offsets(t, dz, dy, dx);
  f32x4 acc[NBO][NBI];
#pragma unroll
  for (int j = 0; j < NBO; ++j)
#pragma unroll
    for (int p = 0; p < NBI; ++p) acc[j][p] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int xch = (v.nx + 31) / 32;
  const long items = (long)v.nb * v.nz * v.ny * xch;
  for (long it = chunk; it < items; it += G) {
    const long row = it / xch;
    const int x0 = (int)(it - row * xch) * 32;
    const int y = (int)(row % v.ny);
    const long bzi = row / v.ny;
    const int z = (int)(bzi % v.nz);
    const int zz = z + dz, yy = y + dy;
    if (zz < 0 || zz >= v.nz || yy < 0 || yy >= v.ny) continue;
    const long srow = (bzi - z + zz) * v.ny + yy;
    const bf16_t* gp = gr + row * (long)v.nx * COUT;
    const bf16_t* ip = in + srow * (long)v.nx * CIN;
    u32x4 a[NBO], b[NBI];
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
      const int x = x0 + 8 * q + i;
      const bool g0 = x < v.nx, g1 = x + 1 < v.nx;
      const bool i0 = g0 && x + dx >= 0 && x + dx < v.nx, i1 = g1 && x + 1 + dx >= 0 && x + 1 + dx < v.nx;
#pragma unroll
      for (int j = 0; j < NBO; ++j) {
        const unsigned v0 = g0 ? gp[(long)x * COUT + 16 * j + r] : 0u;
        const unsigned v1 = g1 ? gp[(long)(x + 1) * COUT + 16 * j + r] : 0u;
        a[j][i / 2] = v0 | (v1 << 16);
      }
#pragma unroll
      for (int p = 0; p < NBI; ++p) {
        const unsigned v0 = i0 ? ip[(long)(x + dx) * CIN + 16 * p + r] : 0u;
        const unsigned v1 = i1 ? ip[(long)(x + 1 + dx) * CIN + 16 * p + r] : 0u;
        b[p][i / 2] = v0 | (v1 << 16);
      }
    }
#pragma unroll
    for (int j = 0; j < NBO; ++j)
#pragma unroll
      for (int p = 0; p < NBI; ++p) acc[j][p] = Mfma<32>::run(a[j], b[p], acc[j][p]);
  }
  float* dst = part + ((long)chunk * 27 + t) * CIN * COUT;
#pragma unroll
  for (int j = 0; j < NBO; ++j)
#pragma unroll
    for (int p = 0; p < NBI; ++p)
      *reinterpret_cast<f32x4*>(dst + (16 * p + r) * COUT + 16 * j + 4 * q) = acc[j][p];
}

// First layer weight gradient: n runs over the taps (two blocks of 16, taps
// >= 27 dropped), the input is bf16(raw / s) as in the forward pass.
__global__ __launch_bounds__(256) void k_wgrad_in1(const float* __restrict__ raw, const float* __restrict__ inv_s,
                                                   const bf16_t* __restrict__ gr, Vol v, int G,
                                                   float* __restrict__ part) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int chunk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (chunk >= G) return;
  int tdz[2], tdy[2], tdx[2];
  tap_offsets(r, tdz[0], tdy[0], tdx[0]);
  tap_offsets(16 + r, tdz[1], tdy[1], tdx[1]);
  const bool tv1 = 16 + r < 27;
  f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
  const int xch = (v.nx + 31) / 32;
  const long items = (long)v.nb * v.nz * v.ny * xch;
  for (long it = chunk; it < items; it += G) {
    const long row = it / xch;
    const int x0 = (int)(it - row * xch) * 32;
    const int y = (int)(row % v.ny);
    const long bzi = row / v.ny;
    const int z = (int)(bzi % v.nz);
    const long b = bzi / v.nz;
    const float sc = inv_s[b];
    const float* smp = raw + b * (long)v.nz * v.ny * v.nx;
    const bf16_t* gp = gr + row * (long)v.nx * 16;
    u32x4 a, bb[2];
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
      const int x = x0 + 8 * q + i;
      const unsigned g0 = x < v.nx ? gp[(long)x * 16 + r] : 0u;
      const unsigned g1 = x + 1 < v.nx ? gp[(long)(x + 1) * 16 + r] : 0u;
      a[i / 2] = g0 | (g1 << 16);
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        float u[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int xo = x + h, zz = z + tdz[p], yy = y + tdy[p], xx = xo + tdx[p];
          const bool ok = (p == 0 || tv1) && xo < v.nx && zz >= 0 && zz < v.nz && yy >= 0 && yy < v.ny && xx >= 0 &&
                          xx < v.nx;
          u[h] = ok ? smp[((long)zz * v.ny + yy) * v.nx + xx] * sc : 0.f;
        }
        bb[p][i / 2] = pack2(u[0], u[1]);
      }
    }
    acc[0] = Mfma<32>::run(a, bb[0], acc[0]);
    acc[1] = Mfma<32>::run(a, bb[1], acc[1]);
  }
  // D[m = co = 4q + i][n = tap = 16p + r] -> part[chunk][tap][co]
  float* dst = part + (long)chunk * 27 * 16;
  *reinterpret_cast<f32x4*>(dst + r * 16 + 4 * q) = acc[0];
  if (tv1) *reinterpret_cast<f32x4*>(dst + (16 + r) * 16 + 4 * q) = acc[1];
}

// out[c] = sum_g part[g][c] (double, fixed order); accumulate into out or set
__global__ void k_reduce_rows(const float* __restrict__ part, int rows, long cols, float* __restrict__ out,
                              double scale) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < cols; c += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < rows; ++g) s += part[(long)g * cols + c];
    out[c] = (float)(s * scale);
  }
}

// per-sample sums of the FP32 stack: part[b][blk]
constexpr int SUM_THREADS = 256;
__global__ __launch_bounds__(SUM_THREADS) void k_sample_sum(const float* __restrict__ raw, long per, int blocks,
                                                            double* __restrict__ part) {
  const int b = blockIdx.x / blocks, k = blockIdx.x % blocks;
  const float* p = raw + (long)b * per;
  double s = 0.0;
  for (long i = (long)k * SUM_THREADS + threadIdx.x; i < per; i += (long)blocks * SUM_THREADS) s += p[i];
  __shared__ double red[SUM_THREADS];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = SUM_THREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// inv_s[b] = n / sum (the inverse of the sample mean; 1 for an all-zero
// sample); one block per sample, fixed summation order
__global__ __launch_bounds__(SUM_THREADS) void k_inv_scale(const double* __restrict__ part, int blocks, long per,
                                                           float* __restrict__ inv_s) {
  const int b = blockIdx.x;
  double s = 0.0;
  for (int k = threadIdx.x; k < blocks; k += SUM_THREADS) s += part[(long)b * blocks + k];
  __shared__ double red[SUM_THREADS];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = SUM_THREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) inv_s[b] = red[0] > 0.0 ? (float)((double)per / red[0]) : 1.f;
}

// Loss and head gradient from the head partials: loss = sum res^2 / n_total,
// grad[head_off + c] = sum of the dWh partials, grad[head_off + 16] = dbh.
__global__ void k_head_reduce(const float* __restrict__ part, int blocks, double inv_total, float* __restrict__ grad,
                              long head_off, float* __restrict__ loss) {
  const int k = threadIdx.x;
  if (k >= HEAD_PART) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += part[(long)b * HEAD_PART + k];
  if (k == 0)
    *loss = (float)(s * inv_total);
  else
    grad[head_off + k - 1] = (float)s;
}

// Packed BF16 MFMA A operand of one 3x3x3 layer from the FP32 master weights
// W[t][ci][co] (ci < cin, co < cout).  For the backward-data pass (flip) the
// layer is the adjoint: W'[t][co][ci] = W[26 - t][ci][co] with the channel
// counts swapped.  Layout: [t][k-step][co-block][lane][F], lane = 16q + r
// holding W[ci = step*KS + F q + i][co = 16 block + r]; for a single input
// channel K runs over the taps ([lane][8], tap = 8q + i, zero for tap >= 27).
__global__ void k_pack(const float* __restrict__ W, int cin, int cout, int flip, bf16_t* __restrict__ out) {
  const int ci_n = flip ? cout : cin, co_n = flip ? cin : cout;  // the packed layer's channels
  const long n = packed_weights(ci_n, co_n);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    float val;
    if (ci_n == 1) {
      const int i = (int)(e % 8), lane = (int)((e / 8) % 64), j = (int)(e / 512);
      const int tap = 8 * (lane >> 4) + i, co = 16 * j + (lane & 15);
      val = tap < 27 ? W[tap * co_n + co] : 0.f;
    } else if (ci_n == 16) {
      // tap pairs: K = 32 = taps 2p (lanes q < 2) and 2p + 1 (q >= 2), 16
      // channels each; [pair][co-block][lane][8], tap 27 is zero
      const int i = (int)(e % 8), lane = (int)((e / 8) % 64), NBO = co_n / 16;
      const int j = (int)((e / 512) % NBO), pr = (int)(e / (512L * NBO));
      const int q = lane >> 4, t = 2 * pr + (q >> 1), ci = 8 * (q & 1) + i, co = 16 * j + (lane & 15);
      val = t >= 27 ? 0.f
                    : (flip ? W[((long)(26 - t) * cin + co) * cout + ci] : W[((long)t * cin + ci) * cout + co]);
    } else {
      const int KS = ci_n == 16 ? 16 : 32, F = KS / 4, NS = ci_n / KS, NBO = co_n / 16;
      long k = e;
      const int i = (int)(k % F);
      k /= F;
      const int lane = (int)(k % 64);
      k /= 64;
      const int j = (int)(k % NBO);
      k /= NBO;
      const int s = (int)(k % NS);
      const int t = (int)(k / NS);
      const int ci = s * KS + F * (lane >> 4) + i, co = 16 * j + (lane & 15);
      val = flip ? W[((long)(26 - t) * cin + co) * cout + ci] : W[((long)t * cin + ci) * cout + co];
    }
    out[e] = f2bf(val);
  }
}

// Adam (bias-corrected step size folded into lr_t; c2 = 1 / (1 - beta2^t))
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ vv, long n, float lr_t, float b1, float b2, float eps, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * vv[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    vv[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi * c2) + eps);
  }
}

int grid_for(long n, int threads, int cap = 4096) {
  long g = (n + threads - 1) / threads;
  return (int)std::max(1L, std::min<long>(g, cap));
}

#ifndef CDN_CONV25
#define CDN_CONV25 1
#endif
template <int CIN, int COUT, bool RELU, bool BIAS, bool UPADD = false, bool POOL = false>
void conv25_launch(hipStream_t st, const bf16_t* in, const bf16_t* wf, const float* bias, bf16_t* out, Vol v,
                   const bf16_t* coarse = nullptr, bf16_t* pooled = nullptr) {
  constexpr int TX = CIN == 64 ? 32 : 64;
  // z-chunks of 16 planes (2 extra planes of loads per chunk), fewer when the
  // grid is already large
  const long tiles = (long)v.nb * ((v.ny + 3) / 4) * ((v.nx + TX - 1) / TX);
  int zchunk = v.nz < 16 ? v.nz : 16;
  if (tiles >= 4096) zchunk = v.nz < 32 ? v.nz : 32;
  const long grid = tiles * ((v.nz + zchunk - 1) / zchunk);
  hipLaunchKernelGGL((k_conv25<CIN, COUT, TX, RELU, BIAS, UPADD, POOL>), dim3((unsigned)grid), dim3(256), 0, st, in,
                     wf, bias, out, v, zchunk, coarse, pooled);
}

#ifndef CDN_CONV_MB
#define CDN_CONV_MB 4
#endif
template <int CIN, int COUT, bool RELU, bool BIAS>
void conv_plain_launch(hipStream_t st, const bf16_t* in, const bf16_t* wf, const float* bias, bf16_t* out, Vol v) {
  // 16-voxel blocks per wave; 2 for COUT = 64 (more would spill at 1024 threads)
  constexpr int MB = COUT == 64 ? 2 : CDN_CONV_MB;
  constexpr int THREADS = 27 * CIN * COUT * 2 > 64 * 1024 ? 1024 : 256;
  const long items = (long)v.nb * v.nz * v.ny * ((v.nx + 16 * MB - 1) / (16 * MB));
  const int waves = THREADS / 64;
  // grid-stride kernel: launch what is resident at once (no tail round)
  static int resident = 0;
  if (resident == 0) {
    int per_cu = 0, dev = 0, cus = 0;
    CDN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv<CIN, COUT, MB, THREADS, RELU, BIAS>,
                                                         THREADS, 0));
    CDN_HIP(hipGetDevice(&dev));
    CDN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    resident = std::max(1, per_cu) * std::max(1, cus);
  }
  const int grid = (int)std::max(1L, std::min<long>((items + waves - 1) / waves, resident));
  hipLaunchKernelGGL((k_conv<CIN, COUT, MB, THREADS, RELU, BIAS>), dim3(grid), dim3(THREADS), 0, st, in, wf, bias,
                     out, v);
}

// 64-channel layers use k_conv (the blocked kernel would spill registers);
// 16 input channels need the blocked kernel (tap-pair operand layout)
template <int CIN, int COUT, bool RELU, bool BIAS>
void conv_launch(hipStream_t st, const bf16_t* in, const bf16_t* wf, const float* bias, bf16_t* out, Vol v) {
  if constexpr (CIN == 16 || CDN_CONV25)
    conv25_launch<CIN, COUT, RELU, BIAS>(st, in, wf, bias, out, v);
  else
    conv_plain_launch<CIN, COUT, RELU, BIAS>(st, in, wf, bias, out, v);
}

template <int CIN, int COUT>
void wgrad_launch(hipStream_t st, const bf16_t* in, const bf16_t* g, Vol v, int G, float* part) {
  hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3((27 * G + 3) / 4), dim3(256), 0, st, in, g, v, G, part);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void net_conv(hipStream_t st, int cin, int cout, bool relu, const bf16_t* in, const bf16_t* wf, const float* bias,
              bf16_t* out, Vol v) {
  const int key = cin * 1000 + cout;
  if (relu) {
    switch (key) {
      case 16032: return conv_launch<16, 32, true, true>(st, in, wf, bias, out, v);
      case 32064: return conv_launch<32, 64, true, true>(st, in, wf, bias, out, v);
      case 64032: return conv_launch<64, 32, true, true>(st, in, wf, bias, out, v);
      case 32016: return conv_launch<32, 16, true, true>(st, in, wf, bias, out, v);
    }
  } else {
    switch (key) {
      case 16032: return conv_launch<16, 32, false, false>(st, in, wf, bias, out, v);
      case 32064: return conv_launch<32, 64, false, false>(st, in, wf, bias, out, v);
      case 64032: return conv_launch<64, 32, false, false>(st, in, wf, bias, out, v);
      case 32016: return conv_launch<32, 16, false, false>(st, in, wf, bias, out, v);
    }
  }
  fail("net_conv: unsupported channel pair");
}

void net_conv_upadd_32_16(hipStream_t st, const bf16_t* coarse, const bf16_t* skip, const bf16_t* wf,
                          const float* bias, bf16_t* out, Vol v) {
  conv25_launch<32, 16, true, true, true>(st, skip, wf, bias, out, v, coarse);
}

void net_conv_pool_16_32(hipStream_t st, const bf16_t* in, const bf16_t* wf, const float* bias, bf16_t* out,
                         bf16_t* pooled, Vol v) {
  if (v.nz % 2 || v.ny % 2 || v.nx % 2) fail("net_conv_pool_16_32: dimensions must be even");
  conv25_launch<16, 32, true, true, false, true>(st, in, wf, bias, out, v, nullptr, pooled);
}

void net_first(hipStream_t st, int mode, const float* raw, const float* inv_s, const bf16_t* wf, const float* bias,
               bf16_t* out16, const bf16_t* d2, const float* hp, float* out32, Vol v) {
  if (v.nz % 2 != 0) fail("net_first: nz must be even");
  // z-chunks of 16 planes (even, so z pairs never straddle chunks), 32 when
  // the grid is already large
  const long tiles = (long)v.nb * ((v.ny + 7) / 8) * ((v.nx + 63) / 64);
  int zchunk = v.nz < 16 ? v.nz : 16;
  if (tiles >= 4096) zchunk = v.nz < 32 ? v.nz : 32;
  const long grid = tiles * ((v.nz + zchunk - 1) / zchunk);
  if (grid > 0x7fffffffL) fail("net_first: volume too large");
  switch (mode) {
    case 0:
      hipLaunchKernelGGL(k_first<0>, dim3(grid), dim3(256), 0, st, raw, inv_s, wf, bias, out16, d2, hp, out32, v,
                         zchunk);
      return;
    case 1:
      hipLaunchKernelGGL(k_first<1>, dim3(grid), dim3(256), 0, st, raw, inv_s, wf, bias, out16, d2, hp, out32, v,
                         zchunk);
      return;
    case 2:
      hipLaunchKernelGGL(k_first<2>, dim3(grid), dim3(256), 0, st, raw, inv_s, wf, bias, out16, d2, hp, out32, v,
                         zchunk);
      return;
    case 3:
      hipLaunchKernelGGL(k_first<3>, dim3(grid), dim3(256), 0, st, raw, inv_s, wf, bias, out16, d2, hp, out32, v,
                         zchunk);
      return;
  }
  fail("net_first: mode");
}

void net_pool(hipStream_t st, int c, const bf16_t* in, bf16_t* out, Vol fine) {
  const long n = (long)fine.nb * (fine.nz / 2) * (fine.ny / 2) * (fine.nx / 2) * (c / 8);
  const int grid = grid_for(n, 256);
  if (c == 16) hipLaunchKernelGGL(k_pool<16>, dim3(grid), dim3(256), 0, st, in, out, fine);
  else if (c == 32) hipLaunchKernelGGL(k_pool<32>, dim3(grid), dim3(256), 0, st, in, out, fine);
  else fail("net_pool: channels");
}

void net_upadd(hipStream_t st, int c, const bf16_t* coarse, const bf16_t* skip, bf16_t* out, Vol fine) {
  const long n = (long)fine.nb * fine.nz * fine.ny * fine.nx * (c / 8);
  const int grid = grid_for(n, 256);
  if (c == 32) hipLaunchKernelGGL(k_upadd<32>, dim3(grid), dim3(256), 0, st, coarse, skip, out, fine);
  else if (c == 16) hipLaunchKernelGGL(k_upadd<16>, dim3(grid), dim3(256), 0, st, coarse, skip, out, fine);
  else fail("net_upadd: channels");
}

int net_head_blocks(Vol v) { return grid_for((long)v.nb * v.nz * v.ny * v.nx, HEAD_THREADS, 1024); }

void net_head(hipStream_t st, int mode, const float* raw, const float* inv_s, const bf16_t* e1, const bf16_t* d2,
              const float* hp, const float* tgt, float* out, float* part, Vol v) {
  const int grid = net_head_blocks(v);
  const float inv_total = 1.f / (float)((double)v.nb * v.nz * v.ny * v.nx);
  if (mode == 0)
    hipLaunchKernelGGL(k_head<0>, dim3(grid), dim3(HEAD_THREADS), 0, st, raw, inv_s, e1, d2, hp, tgt, out, part, v,
                       inv_total);
  else if (mode == 1)
    hipLaunchKernelGGL(k_head<1>, dim3(grid), dim3(HEAD_THREADS), 0, st, raw, inv_s, e1, d2, hp, tgt, out, part, v,
                       inv_total);
  else
    hipLaunchKernelGGL(k_head<2>, dim3(grid), dim3(HEAD_THREADS), 0, st, raw, inv_s, e1, d2, hp, tgt, out, part, v,
                       inv_total);
}

void net_head_reduce(hipStream_t st, const float* part, int blocks, double inv_total, float* grad, long head_off,
                     float* loss) {
  hipLaunchKernelGGL(k_head_reduce, dim3(1), dim3(64), 0, st, part, blocks, inv_total, grad, head_off, loss);
}

void net_grad_enc(hipStream_t st, int c, const bf16_t* e, const bf16_t* skip, const float* dout, const float* wh,
                  const bf16_t* dP, bf16_t* dpre, Vol v) {
  const int grid = grid_for((long)v.nb * v.nz * v.ny * v.nx * (c / 8), 256);
  if (c == 16) hipLaunchKernelGGL(k_grad_enc<16>, dim3(grid), dim3(256), 0, st, e, skip, dout, wh, dP, dpre, v);
  else if (c == 32) hipLaunchKernelGGL(k_grad_enc<32>, dim3(grid), dim3(256), 0, st, e, skip, dout, wh, dP, dpre, v);
  else if (c == 64) hipLaunchKernelGGL(k_grad_enc<64>, dim3(grid), dim3(256), 0, st, e, skip, dout, wh, dP, dpre, v);
  else fail("net_grad_enc: channels");
}

void net_grad_dec(hipStream_t st, int c, const bf16_t* d, const bf16_t* fine, const float* dout, const float* wh,
                  bf16_t* dpre, Vol coarse) {
  const int grid = grid_for((long)coarse.nb * coarse.nz * coarse.ny * coarse.nx * (c / 8), 256);
  if (c == 16) hipLaunchKernelGGL(k_grad_dec<16>, dim3(grid), dim3(256), 0, st, d, fine, dout, wh, dpre, coarse);
  else if (c == 32) hipLaunchKernelGGL(k_grad_dec<32>, dim3(grid), dim3(256), 0, st, d, fine, dout, wh, dpre, coarse);
  else fail("net_grad_dec: channels");
}

void net_chan_sum(hipStream_t st, int c, const bf16_t* t, long nvox, int blocks, float* part) {
  if (c == 16) hipLaunchKernelGGL(k_chan_sum<16>, dim3(blocks), dim3(256), 0, st, t, nvox, part);
  else if (c == 32) hipLaunchKernelGGL(k_chan_sum<32>, dim3(blocks), dim3(256), 0, st, t, nvox, part);
  else if (c == 64) hipLaunchKernelGGL(k_chan_sum<64>, dim3(blocks), dim3(256), 0, st, t, nvox, part);
  else fail("net_chan_sum: channels");
}

void net_wgrad(hipStream_t st, int cin, int cout, const bf16_t* in, const bf16_t* g, Vol v, int G, float* part) {
  switch (cin * 1000 + cout) {
    case 16032: return wgrad_launch<16, 32>(st, in, g, v, G, part);
    case 32064: return wgrad_launch<32, 64>(st, in, g, v, G, part);
    case 64032: return wgrad_launch<64, 32>(st, in, g, v, G, part);
    case 32016: return wgrad_launch<32, 16>(st, in, g, v, G, part);
  }
  fail("net_wgrad: unsupported channel pair");
}

void net_wgrad_in1(hipStream_t st, const float* raw, const float* inv_s, const bf16_t* g, Vol v, int G, float* part) {
  hipLaunchKernelGGL(k_wgrad_in1, dim3((G + 3) / 4), dim3(256), 0, st, raw, inv_s, g, v, G, part);
}

void net_reduce_rows(hipStream_t st, const float* part, int rows, long cols, float* out) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(grid_for(cols, 256, 1024)), dim3(256), 0, st, part, rows, cols, out, 1.0);
}

int net_sum_blocks(int nb) { return std::max(64, 2048 / std::max(nb, 1)); }

void net_inv_scale(hipStream_t st, const float* raw, Vol v, double* part, float* inv_s) {
  const long per = (long)v.nz * v.ny * v.nx;
  const int blocks = net_sum_blocks(v.nb);
  hipLaunchKernelGGL(k_sample_sum, dim3(v.nb * blocks), dim3(SUM_THREADS), 0, st, raw, per, blocks, part);
  hipLaunchKernelGGL(k_inv_scale, dim3(v.nb), dim3(SUM_THREADS), 0, st, part, blocks, per, inv_s);
}

long net_packed_weights(int cin, int cout) { return packed_weights(cin, cout); }

void net_pack(hipStream_t st, const float* W, int cin, int cout, int flip, bf16_t* out) {
  const long n = packed_weights(flip ? cout : cin, flip ? cin : cout);
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, W, cin, cout, flip, out);
}

void net_adam(hipStream_t st, float* p, const float* g, float* m, float* v, long n, float lr_t, float b1, float b2,
              float eps, float c2) {
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, p, g, m, v, n, lr_t, b1, b2, eps, c2);
}

}  // namespace cdn
