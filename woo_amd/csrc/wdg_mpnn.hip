// Learned coarse step: a message-passing network on the contact graph, trained and run by
// hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Network (hidden width H = 64):
//   h0 = relu(X W_enc + b_enc)                       X: node inputs, 32 columns (5 used)
//   for l in 0, 1:
//     m_e   = relu([h_recv(e), h_send(e), F_e] W_m,l + b_m,l)   F_e: edge inputs, 32 columns (9 used)
//     agg_i = sum over edges e received by i of m_e
//     u_i   = relu([h_i, agg_i] W_u,l + b_u,l);  h_i <- h_i + u_i
//   a_i = h_i W_dec + b_dec                          (3 outputs: acceleration / g)
// Storage: activations that feed a GEMM are BF16 (h, m, agg, u); GEMMs accumulate in FP32;
// the residual sum is formed in FP32 and rounded once; parameters are FP32 master copies with
// BF16 images (transposed for the forward GEMM operand, row-major for the input-gradient one).
//
// Graph: edges in CSR by receiver (rowOff, recv, send), both directions of every pair, and
// rev[e] = the opposite edge; every sum over edges follows this order (deterministic).
//
// GEMM tiling for 64-wide wavefronts: one wavefront computes a 16-row x 64-column block as
// four 16x16 MFMA accumulators over K in steps of 32; lane l supplies row (l mod 16), 8
// consecutive k from 8*(l/16) as one 16-byte load, gathered straight from the node and edge
// arrays (the concatenated GEMM inputs are never stored). The weight-gradient GEMM reduces
// over rows: each workgroup stages 32-row chunks of the input and of the output gradient in
// LDS transposed, so that the MFMA operands are again 16-byte reads, and writes a partial;
// a second kernel sums the partials in a fixed order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "wdg_internal.hpp"

namespace wdg {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef uint16_t bf16;

constexpr int H = 64;     // hidden width
constexpr int KX = 32;    // node input columns (NIN used)
constexpr int KE = 32;    // edge input columns (NEI used)
constexpr int NIN = 7, NEI = 15;
constexpr int KM = 2 * H + KE;  // message GEMM inner dimension (160)
constexpr int KU = 2 * H;       // update GEMM inner dimension (128)
constexpr int NOUT = 3;
// input scales: lengths by R0, velocities by V0 (outputs are accelerations in units of MPNN_A0)
// overlaps (fractions of the contact distance, 1e-3 .. 1e-2) are scaled by MPNN_OV so that BF16
// resolves them, and clipped to +-MPNN_OV_CLIP as inputs of their own (gaps of spheres that are
// far apart would otherwise dominate the first layers); the force inputs use them unclipped
constexpr double MPNN_R0 = 2.5e-3, MPNN_V0 = 1.0, MPNN_OV = 100.0, MPNN_OV_CLIP = 2.0;
// unit of the outputs and targets: 100 g, the order of contact accelerations, so that outputs are
// of order one
constexpr double MPNN_A0 = 100.0 * 9.81;

// Inputs of edge e (receiver i, sender j), all scaled, computed in double:
//   0-2   (x_j - x_i) / (r_i + r_j)
//   3     overlap ov = (1 - d / (r_i + r_j)) OV, clipped to +-OV_CLIP
//   4-6   (v_j - v_i) / V0
//   7, 8  r_i / R0, r_j / R0
//   9-11  -n ov+^(3/2) sqrt(a b) (a + b) / a^3            (n = (x_j - x_i) / d, a = r_i / R0,
//                                                          b = r_j / R0, ov+ = max(ov, 0)):
//         the direction and radius scaling of a Hertz force on i, divided by i's mass
//   12-14 n vN ov+^(1/4) (a + b)^(1/4) sqrt(mb sqrt(Rb)) / a^3   (vN = (v_j - v_i).n / V0,
//         mb = a^3 b^3 / (a^3 + b^3), Rb = a b / (a + b)): the same for the viscous term
// Inputs of node i (floor at height zf, h = z_i - zf):
//   0-2   v_i / V0;  3  r_i / R0;  4  clip((1 - h / r_i) OV, +-OV_CLIP)
//   5     ovf+^(3/2) / a;  6  (v_i)_z / V0 ovf+^(1/4) / a   (ovf+ = max((1 - h / r_i) OV, 0))
// The same expressions, in the same order, build host graphs (woo_amd/surrogate.py).
__host__ __device__ inline void edge_inputs(double xi, double yi, double zi, double xj, double yj, double zj,
                                            double vxi, double vyi, double vzi, double vxj, double vyj,
                                            double vzj, double ri, double rj, float* f) {
  const double rs = ri + rj;
  const double dx = (xj - xi) / rs, dy = (yj - yi) / rs, dz = (zj - zi) / rs;
  const double dn = sqrt((dx * dx + dy * dy) + dz * dz);
  const double ov = (1. - dn) * MPNN_OV;
  const double wx = (vxj - vxi) / MPNN_V0, wy = (vyj - vyi) / MPNN_V0, wz = (vzj - vzi) / MPNN_V0;
  const double a = ri / MPNN_R0, b = rj / MPNN_R0;
  const double nx = dn > 0 ? dx / dn : 0., ny = dn > 0 ? dy / dn : 0., nz = dn > 0 ? dz / dn : 0.;
  const double ovp = ov > 0 ? ov : 0.;
  const double a3 = a * a * a, b3 = b * b * b;
  const double hz = ovp * sqrt(ovp) * (sqrt(a * b) * (a + b) / a3);
  const double mb = a3 * b3 / (a3 + b3), Rb = a * b / (a + b);
  const double vn = (wx * nx + wy * ny) + wz * nz;
  const double dp = vn * sqrt(sqrt(ovp)) * (sqrt(sqrt(a + b)) * sqrt(mb * sqrt(Rb)) / a3);
  const double ovc = fmin(fmax(ov, -MPNN_OV_CLIP), MPNN_OV_CLIP);
  const double v[NEI] = {dx, dy, dz, ovc, wx, wy, wz, a, b, -nx * hz, -ny * hz, -nz * hz, nx * dp, ny * dp, nz * dp};
  for (int q = 0; q < NEI; q++) f[q] = (float)v[q];
}
__host__ __device__ inline void node_inputs(double vx, double vy, double vz, double r, double h, float* f) {
  const double a = r / MPNN_R0;
  const double ov = (1. - h / r) * MPNN_OV;
  const double ovp = ov > 0 ? ov : 0.;
  const double v[NIN] = {vx / MPNN_V0, vy / MPNN_V0, vz / MPNN_V0, a, fmin(fmax(ov, -MPNN_OV_CLIP), MPNN_OV_CLIP),
                         ovp * sqrt(ovp) / a, vz / MPNN_V0 * sqrt(sqrt(ovp)) / a};
  for (int q = 0; q < NIN; q++) f[q] = (float)v[q];
}

// round to nearest even (finite inputs)
__host__ __device__ inline bf16 f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// ------------------------------------------------------------------ parameter layout (floats)
struct PLayout {
  size_t encW, encB, mW[2], mB[2], uW[2], uB[2], decW, decB, total;
  PLayout() {
    size_t o = 0;
    encW = o; o += (size_t)KX * H;
    encB = o; o += H;
    for (int l = 0; l < 2; l++) {
      mW[l] = o; o += (size_t)KM * H;
      mB[l] = o; o += H;
      uW[l] = o; o += (size_t)KU * H;
      uB[l] = o; o += H;
    }
    decW = o; o += (size_t)H * NOUT;
    decB = o; o += NOUT;
    total = o;
  }
};

// ------------------------------------------------------------------ GEMM input rows
enum { A_PLAIN = 0, A_EDGE = 1, A_UPD = 2 };
struct ARows {
  const bf16* x;      // A_PLAIN: [M x K] row-major
  const bf16* h;      // A_EDGE / A_UPD: node activations [N x H]
  const bf16* ef;     // A_EDGE: edge inputs [E x KE]
  const bf16* agg;    // A_UPD: aggregated messages [N x H]
  const int32_t *recv, *send;  // A_EDGE
  int K;
};
// 8 consecutive inputs of row r starting at column k (k multiple of 8)
template <int MODE>
__device__ __forceinline__ uint4 load8(const ARows& a, int64_t r, int k) {
  const bf16* p;
  if (MODE == A_PLAIN) p = a.x + r * a.K + k;
  else if (MODE == A_EDGE) {
    if (k < H) p = a.h + (int64_t)a.recv[r] * H + k;
    else if (k < 2 * H) p = a.h + (int64_t)a.send[r] * H + (k - H);
    else p = a.ef + r * KE + (k - 2 * H);
  } else {
    if (k < H) p = a.h + r * H + k;
    else p = a.agg + r * H + (k - H);
  }
  return *reinterpret_cast<const uint4*>(p);
}
__device__ __forceinline__ bf16x8 as_bf8(uint4 v) {
  bf16x8 r;
  memcpy(&r, &v, 16);
  return r;
}

// ------------------------------------------------------------------ forward GEMM
// Y = relu(A W + b) over M rows, 64 columns. Out: y (BF16, M x 64). With RESID (update
// layer): u = relu(.) to y, h_new = round(h + u) to hOut.
// The workgroup stages the BF16 weight image in LDS once; each wavefront then walks 16-row
// tiles of A (all K/32 row loads of a tile issued before its MFMAs), and writes each output
// tile through its own LDS tile as 16-byte row stores.
template <int MODE>
constexpr int k_of() { return MODE == A_PLAIN ? KX : (MODE == A_EDGE ? KM : KU); }
template <int MODE, bool RESID>
__global__ __launch_bounds__(256) void k_mm_fwd(ARows a, int64_t M, const bf16* __restrict__ Wt,
                                                 const float* __restrict__ bias, bf16* y, bf16* hOut) {
  constexpr int K = k_of<MODE>(), KS = K / 32, WP = K + 8;  // padded LDS rows (bank spread)
  __shared__ bf16 Ws[H * WP];
  __shared__ bf16 tile[4][16][H + 8];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // weight image [64 x K] into LDS, 16 bytes per thread and step
  for (int i = threadIdx.x; i < H * K / 8; i += blockDim.x) {
    const int col = i / (K / 8), k8 = (i % (K / 8)) * 8;
    *reinterpret_cast<uint4*>(&Ws[col * WP + k8]) = *reinterpret_cast<const uint4*>(Wt + (int64_t)col * K + k8);
  }
  __syncthreads();
  const int kq = 8 * (lane >> 4);
  float bv[4];
#pragma unroll
  for (int cb = 0; cb < 4; cb++) bv[cb] = bias[cb * 16 + (lane & 15)];
  const int64_t ntiles = (M + 15) / 16;
  for (int64_t t = (int64_t)blockIdx.x * 4 + wv; t < ntiles; t += (int64_t)gridDim.x * 4) {
    const int64_t r0 = t * 16;
    const int64_t r = r0 + (lane & 15);
    const bool live = r < M;
    uint4 av[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) av[ks] = live ? load8<MODE>(a, r, ks * 32 + kq) : make_uint4(0, 0, 0, 0);
    floatx4 acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const bf16x8 A = as_bf8(av[ks]);
#pragma unroll
      for (int cb = 0; cb < 4; cb++) {
        const int col = cb * 16 + (lane & 15);
        const bf16x8 B = as_bf8(*reinterpret_cast<const uint4*>(&Ws[col * WP + ks * 32 + kq]));
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[cb], 0, 0, 0);
      }
    }
    // C layout: row 4*(lane/16)+q, column lane%16 of each block
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
      const int col = cb * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; q++) tile[wv][4 * (lane >> 4) + q][col] = f2bf(fmaxf(acc[cb][q] + bv[cb], 0.f));
    }
    wave_lds_sync();
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int row = half * 8 + (lane >> 3), c0 = (lane & 7) * 8;
      const int64_t rr = r0 + row;
      if (rr >= M) continue;
      const uint4 v = *reinterpret_cast<const uint4*>(&tile[wv][row][c0]);
      *reinterpret_cast<uint4*>(y + rr * H + c0) = v;
      if (RESID) {
        const uint4 hv = *reinterpret_cast<const uint4*>(a.h + rr * H + c0);
        const uint32_t uw[4] = {v.x, v.y, v.z, v.w}, hw[4] = {hv.x, hv.y, hv.z, hv.w};
        uint32_t o[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const bf16 lo = f2bf(bf2f((bf16)(hw[q] & 0xFFFFu)) + bf2f((bf16)(uw[q] & 0xFFFFu)));
          const bf16 hi = f2bf(bf2f((bf16)(hw[q] >> 16)) + bf2f((bf16)(uw[q] >> 16)));
          o[q] = (uint32_t)lo | ((uint32_t)hi << 16);
        }
        *reinterpret_cast<uint4*>(hOut + rr * H + c0) = make_uint4(o[0], o[1], o[2], o[3]);
      }
    }
    // the tile is written again by the next iteration only after every lane has read it
    wave_lds_sync();
  }
}
// workgroups for a forward GEMM over M rows: enough wavefronts to fill the chip, each walking a
// few tiles so that the weight staging is amortised
static unsigned fwd_blocks(int64_t M) {
  const int64_t tiles = (M + 15) / 16;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 1024));
}

// ------------------------------------------------------------------ input-gradient GEMM
// dX[:, c] = sum_k dY[:, k] W[c, k] for output columns c < KOUT (W row-major [K x 64], BF16);
// dY BF16 [M x 64]; dX FP32 [M x KOUT]
template <int KOUT>
__global__ __launch_bounds__(256) void k_mm_dx(const bf16* __restrict__ dY, int64_t M, const bf16* __restrict__ W,
                                                float* dX) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = ((int64_t)blockIdx.x * 4 + wv) * 16;
  if (r0 >= M) return;
  const int64_t r = r0 + (lane & 15);
  const bool live = r < M;
  const int kq = 8 * (lane >> 4);
  constexpr int NB = KOUT / 16;
  floatx4 acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; cb++) acc[cb] = floatx4{0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < H; k0 += 32) {
    const uint4 av = live ? *reinterpret_cast<const uint4*>(dY + r * H + k0 + kq) : make_uint4(0, 0, 0, 0);
    const bf16x8 A = as_bf8(av);
#pragma unroll
    for (int cb = 0; cb < NB; cb++) {
      const int col = cb * 16 + (lane & 15);
      const bf16x8 B = as_bf8(*reinterpret_cast<const uint4*>(W + (int64_t)col * H + k0 + kq));
      acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[cb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int col = cb * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int64_t rr = r0 + 4 * (lane >> 4) + q;
      if (rr < M) dX[rr * KOUT + col] = acc[cb][q];
    }
  }
}

// ------------------------------------------------------------------ weight-gradient GEMM
// partial[blk][c][o] = sum over the workgroup's rows m of A[m][c] dY[m][o]; partialB[blk][o]
// = sum of dY[m][o] (bias gradient). Rows in chunks of 32, staged transposed in LDS.
constexpr int DW_CHUNKS = 8;  // 32-row chunks per workgroup
template <int MODE, int KP>
__global__ __launch_bounds__(256) void k_mm_dw(ARows a, int64_t M, const bf16* __restrict__ dY, float* partial,
                                                float* partialB) {
  __shared__ __attribute__((aligned(16))) bf16 At[KP][32 + 8];
  __shared__ __attribute__((aligned(16))) bf16 Dt[H][32 + 8];
  __shared__ float bsum[4][H];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int TILES = (KP / 16) * (H / 16);
  constexpr int TPW = (TILES + 3) / 4;  // output tiles per wavefront
  floatx4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) acc[t] = floatx4{0, 0, 0, 0};
  float bacc = 0.f;  // bias partial of column lane, rows of this wavefront's share
  const int64_t rowBase = (int64_t)blockIdx.x * DW_CHUNKS * 32;
  for (int ch = 0; ch < DW_CHUNKS; ch++) {
    const int64_t m0 = rowBase + ch * 32;
    if (m0 >= M) break;
    // stage: 32 rows x KP inputs and 32 rows x 64 gradients, transposed
    for (int idx = threadIdx.x; idx < 32 * (KP / 8); idx += 256) {
      const int rr = idx / (KP / 8), k8 = (idx % (KP / 8)) * 8;
      const int64_t m = m0 + rr;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (m < M) v = load8<MODE>(a, m, k8);
      bf16 e[8];
      memcpy(e, &v, 16);
#pragma unroll
      for (int q = 0; q < 8; q++) At[k8 + q][rr] = e[q];
    }
    for (int idx = threadIdx.x; idx < 32 * (H / 8); idx += 256) {
      const int rr = idx / (H / 8), k8 = (idx % (H / 8)) * 8;
      const int64_t m = m0 + rr;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (m < M) v = *reinterpret_cast<const uint4*>(dY + m * H + k8);
      bf16 e[8];
      memcpy(e, &v, 16);
#pragma unroll
      for (int q = 0; q < 8; q++) Dt[k8 + q][rr] = e[q];
    }
    __syncthreads();
    // bias: wavefront wv adds rows 8*wv .. 8*wv+7 of column lane
#pragma unroll
    for (int q = 0; q < 8; q++) bacc += bf2f(Dt[lane][8 * wv + q]);
#pragma unroll
    for (int t = 0; t < TPW; t++) {
      const int tile = wv * TPW + t;
      if (tile < TILES) {
        const int cb = tile / (H / 16), ob = tile % (H / 16);
        const int kq = 8 * (lane >> 4);
        const bf16x8 A = as_bf8(*reinterpret_cast<const uint4*>(&At[cb * 16 + (lane & 15)][kq]));
        const bf16x8 B = as_bf8(*reinterpret_cast<const uint4*>(&Dt[ob * 16 + (lane & 15)][kq]));
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* P = partial + (int64_t)blockIdx.x * KP * H;
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int tile = wv * TPW + t;
    if (tile < TILES) {
      const int cb = tile / (H / 16), ob = tile % (H / 16);
#pragma unroll
      for (int q = 0; q < 4; q++) P[(cb * 16 + 4 * (lane >> 4) + q) * H + ob * 16 + (lane & 15)] = acc[t][q];
    }
  }
  bsum[wv][lane] = bacc;
  __syncthreads();
  if (threadIdx.x < H)
    partialB[(int64_t)blockIdx.x * H + threadIdx.x] =
        ((bsum[0][threadIdx.x] + bsum[1][threadIdx.x]) + bsum[2][threadIdx.x]) + bsum[3][threadIdx.x];
}
// grad[i] = sum over blocks of partial[blk][i], in block order
__global__ void k_sum_partials(const float* __restrict__ partial, int64_t nblk, int64_t n, float* grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int64_t b = 0; b < nblk; b++) s += partial[b * n + i];
  grad[i] = s;
}
// first level of a two-level sum for many blocks: segment y of len blocks, in block order
__global__ void k_sum_seg(const float* __restrict__ partial, int64_t nblk, int64_t n, int64_t len, float* seg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b0 = (int64_t)blockIdx.y * len, b1 = b0 + len < nblk ? b0 + len : nblk;
  float s = 0.f;
  for (int64_t b = b0; b < b1; b++) s += partial[b * n + i];
  seg[(int64_t)blockIdx.y * n + i] = s;
}

// ------------------------------------------------------------------ graph sums
// agg_i = sum of m_e over edges received by i, in edge order (eight threads per node, eight
// features each, 16-byte loads and stores)
__global__ __launch_bounds__(256) void k_agg(int64_t N, const int32_t* __restrict__ rowOff,
                                              const bf16* __restrict__ m, bf16* agg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t >> 3;
  const int f0 = (int)(t & 7) * 8;
  if (i >= N) return;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int32_t e1 = rowOff[i + 1];
  for (int32_t e = rowOff[i]; e < e1; e++) {
    const uint4 v = *reinterpret_cast<const uint4*>(m + (int64_t)e * H + f0);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      s[2 * q] += bf2f((bf16)(w[q] & 0xFFFFu));
      s[2 * q + 1] += bf2f((bf16)(w[q] >> 16));
    }
  }
  uint32_t o[4];
#pragma unroll
  for (int q = 0; q < 4; q++) o[q] = (uint32_t)f2bf(s[2 * q]) | ((uint32_t)f2bf(s[2 * q + 1]) << 16);
  *reinterpret_cast<uint4*>(agg + i * H + f0) = make_uint4(o[0], o[1], o[2], o[3]);
}
// dY_m,e = dagg_recv(e) * (m_e > 0), BF16
__global__ void k_edge_dy(int64_t E, const int32_t* __restrict__ recv, const float* __restrict__ dU,
                          const bf16* __restrict__ m, bf16* dYm) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * H) return;
  const int64_t e = t / H;
  const int f = (int)(t % H);
  const float g = dU[(int64_t)recv[e] * KU + H + f];  // the agg part of the update input gradient
  dYm[t] = f2bf(bf2f(m[t]) > 0.f ? g : 0.f);
}
// dh_i = dhOut_i + dU_i[:H] + sum over edges e received by i of dZ_e[:H] + dZ_rev(e)[H:2H]
__global__ __launch_bounds__(256) void k_node_grad(int64_t N, const int32_t* __restrict__ rowOff,
                                                    const int32_t* __restrict__ rev, const float* __restrict__ dhOut,
                                                    const float* __restrict__ dU, const float* __restrict__ dZ,
                                                    float* dh) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int f = threadIdx.x & 63;
  if (i >= N) return;
  float s = dhOut[i * H + f] + dU[i * KU + f];
  for (int32_t e = rowOff[i]; e < rowOff[i + 1]; e++) {
    s += dZ[(int64_t)e * KU + f];
    s += dZ[(int64_t)rev[e] * KU + H + f];
  }
  dh[i * H + f] = s;
}
// dY = g * (act > 0), BF16 (backward of a relu whose output is act)
__global__ void k_relu_bwd(int64_t n, const float* __restrict__ g, const bf16* __restrict__ act, bf16* dY) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dY[t] = f2bf(bf2f(act[t]) > 0.f ? g[t] : 0.f);
}

// ------------------------------------------------------------------ decoder and loss
// a_i = h_i W_dec + b_dec (FP32), summed over k in order; one thread per node, the row read
// as eight 16-byte loads, the weights from LDS
__global__ __launch_bounds__(256) void k_dec(int64_t N, const bf16* __restrict__ h, const float* __restrict__ W,
                                              const float* __restrict__ b, float* out) {
  __shared__ float Ws[H * NOUT];
  for (int q = threadIdx.x; q < H * NOUT; q += blockDim.x) Ws[q] = W[q];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint4 hv[H / 8];
#pragma unroll
  for (int c8 = 0; c8 < H / 8; c8++) hv[c8] = *reinterpret_cast<const uint4*>(h + i * H + c8 * 8);
  float o[NOUT] = {b[0], b[1], b[2]};
#pragma unroll
  for (int c8 = 0; c8 < H / 8; c8++) {
    const uint32_t w[4] = {hv[c8].x, hv[c8].y, hv[c8].z, hv[c8].w};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = c8 * 8 + q;
      const float x = bf2f((bf16)((q & 1) ? (w[q >> 1] >> 16) : (w[q >> 1] & 0xFFFFu)));
#pragma unroll
      for (int c = 0; c < NOUT; c++) o[c] += x * Ws[k * NOUT + c];
    }
  }
#pragma unroll
  for (int c = 0; c < NOUT; c++) out[i * NOUT + c] = o[c];
}
// loss = mean over nodes of |a - t|^2; dA = 2 (a - t) / N; per-block partial loss
// A node is in contact when its floor overlap input or the overlap input of one of its edges
// is positive. Nodes not in contact move in free flight (the network's output is not used for
// them), so they carry no loss.
__global__ __launch_bounds__(256) void k_contact(int64_t N, const int32_t* __restrict__ rowOff,
                                                  const bf16* __restrict__ X, const bf16* __restrict__ EF,
                                                  uint8_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  bool c = bf2f(X[i * KX + 4]) > 0.f;
  for (int32_t e = rowOff[i]; e < rowOff[i + 1] && !c; e++) c = bf2f(EF[(int64_t)e * KE + 3]) > 0.f;
  flag[i] = c ? 1 : 0;
}
__global__ __launch_bounds__(256) void k_loss(int64_t N, const float* __restrict__ a, const float* __restrict__ t,
                                               const uint8_t* __restrict__ flag, float* dA, float* partLoss) {
  __shared__ float sm[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < N)
    for (int c = 0; c < NOUT; c++) {
      const float d = flag[i] ? a[i * NOUT + c] - t[i * NOUT + c] : 0.f;
      l += d * d;
      dA[i * NOUT + c] = 2.f * d / (float)N;
    }
  sm[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partLoss[blockIdx.x] = sm[0];
}
// decoder backward: dh_i = dA_i W^T (FP32), per-block partials of dW (H x 3) and db (3)
__global__ __launch_bounds__(256) void k_dec_bwd(int64_t N, const bf16* __restrict__ h, const float* __restrict__ W,
                                                  const float* __restrict__ dA, float* dh, float* partW) {
  __shared__ float sm[256][NOUT + 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < N;
  float g[NOUT] = {0, 0, 0};
  if (live)
    for (int c = 0; c < NOUT; c++) g[c] = dA[i * NOUT + c];
  if (live)
    for (int k = 0; k < H; k++) dh[i * H + k] = (g[0] * W[k * NOUT + 0] + g[1] * W[k * NOUT + 1]) + g[2] * W[k * NOUT + 2];
  // dW[k][c] = sum_i h[i][k] g[c]: one pass per k over the block
  for (int k = 0; k < H + 1; k++) {
    const float x = k < H ? (live ? bf2f(h[i * H + k]) : 0.f) : (live ? 1.f : 0.f);  // k == H: bias
    for (int c = 0; c < NOUT; c++) sm[threadIdx.x][c] = x * g[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        for (int c = 0; c < NOUT; c++) sm[threadIdx.x][c] += sm[threadIdx.x + s][c];
      __syncthreads();
    }
    if (threadIdx.x < NOUT) partW[(int64_t)blockIdx.x * (H + 1) * NOUT + k * NOUT + threadIdx.x] = sm[0][threadIdx.x];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ optimizer and images
__global__ void k_adam(int64_t n, float* p, const float* __restrict__ g, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// BF16 images of a K x 64 weight: transposed [64 x K] and row-major [K x 64]
__global__ void k_images(const float* __restrict__ W, int K, bf16* Wt, bf16* Wr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * H) return;
  const int k = t / H, o = t % H;
  const bf16 b = f2bf(W[t]);
  Wr[t] = b;
  Wt[o * K + k] = b;
}

// ------------------------------------------------------------------ state
struct Mpnn {
  int device = 0;
  hipStream_t stream = nullptr;
  PLayout L;
  DBuf<float> P, G, Am, Av;   // parameters, gradients, Adam moments
  DBuf<bf16> img;             // BF16 images (per weight: transposed then row-major)
  size_t imgT[2 + 4], imgR[2 + 4];  // offsets: enc, m0, u0, m1, u1
  int64_t adamStep = 0;
  // graph
  int64_t N = 0, E = 0;
  DBuf<int32_t> rowOff, recv, send, rev;
  DBuf<bf16> X, EF;
  DBuf<float> T;  // targets
  DBuf<int32_t> deg, slotSide, edgeOfSlot;
  // graph from an engine context: the host row of each node in the output (spheres in user-id
  // order); empty for graphs given by the caller
  std::vector<int64_t> outRow;
  Ctx* fromCtx = nullptr;  // the engine context of the current graph (storage order of its nodes)
  DBuf<uint32_t> degScan;
  DBuf<uint32_t> nodeUid;  // user id of each node of a graph from an engine context
  uint64_t uidCtx = 0;  // the context (serial), refresh and upload count nodeUid was taken at
  int64_t uidRefresh = -1, uidN = -1;
  uint64_t uidUploads = 0;
  // activations
  DBuf<bf16> h[3], m[2], agg[2], u[2];
  DBuf<float> out, dA, dhA, dhB, dU, dZ, part, partB, partLoss, partSeg;
  DBuf<bf16> dY, dYe;
  DBuf<uint8_t> cflag;  // node in contact (else in free flight)
  hipEvent_t done = nullptr;  // orders the network's stream and an engine context's stream
  // graphs kept on the device (training sets): inputs, structure and targets
  struct Kept {
    int64_t N = -1, E = 0;
    DBuf<int32_t> rowOff, recv, send, rev;
    DBuf<bf16> X, EF;
    DBuf<float> T;
    void release() {
      rowOff.release(); recv.release(); send.release(); rev.release(); X.release(); EF.release(); T.release();
      N = -1;
    }
  };
  std::vector<Kept> kept;
};

// output row of each node: its own index for graphs given by the caller, the rank of its user
// id among the spheres for a graph from an engine context
static void out_rows(Mpnn& s) {
  if (!s.fromCtx || !s.outRow.empty() || s.N == 0) return;
  // the user ids of the nodes as they were when the graph was built (the context may have
  // re-sorted its particles since)
  std::vector<uint32_t> uid((size_t)s.N);
  WDG_HIP(hipMemcpyAsync(uid.data(), s.nodeUid.p, s.N * 4, hipMemcpyDeviceToHost, s.stream));
  WDG_HIP(hipStreamSynchronize(s.stream));
  std::vector<int64_t> idx((size_t)s.N);
  for (int64_t t = 0; t < s.N; t++) idx[t] = t;
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return uid[a] < uid[b]; });
  s.outRow.assign((size_t)s.N, 0);
  for (int64_t r = 0; r < s.N; r++) s.outRow[idx[r]] = r;
}

template <class T>
static void copy_buf(DBuf<T>& dst, const DBuf<T>& src, size_t n, hipStream_t st) {
  dst.ensure(std::max<size_t>(n, 1));
  if (n) WDG_HIP(hipMemcpyAsync(dst.p, src.p, n * sizeof(T), hipMemcpyDeviceToDevice, st));
}

static void images(Mpnn& s) {
  const int Ks[5] = {KX, KM, KU, KM, KU};
  const size_t src[5] = {s.L.encW, s.L.mW[0], s.L.uW[0], s.L.mW[1], s.L.uW[1]};
  size_t o = 0;
  for (int q = 0; q < 5; q++) {
    s.imgT[q] = o; o += (size_t)Ks[q] * H;
    s.imgR[q] = o; o += (size_t)Ks[q] * H;
  }
  s.img.ensure(o);
  for (int q = 0; q < 5; q++) {
    hipLaunchKernelGGL(k_images, dim3(nblocks((int64_t)Ks[q] * H, 256)), dim3(256), 0, s.stream, s.P.p + src[q], Ks[q],
                       s.img.p + s.imgT[q], s.img.p + s.imgR[q]);
    WDG_LAUNCH_CHECK();
  }
}

// sums of per-block partials in a fixed order: directly for few blocks, else over SUM_SEGS
// segments of consecutive blocks and then over the segments
constexpr int64_t SUM_SEGS = 64;
static void sum_partials(Mpnn& s, const float* partial, int64_t nblk, int64_t n, float* out) {
  if (nblk <= 2 * SUM_SEGS) {
    hipLaunchKernelGGL(k_sum_partials, dim3(nblocks(n, 256)), dim3(256), 0, s.stream, partial, nblk, n, out);
    WDG_LAUNCH_CHECK();
    return;
  }
  const int64_t len = (nblk + SUM_SEGS - 1) / SUM_SEGS;
  hipLaunchKernelGGL(k_sum_seg, dim3(nblocks(n, 256), (unsigned)SUM_SEGS), dim3(256), 0, s.stream, partial, nblk, n,
                     len, s.partSeg.p);
  WDG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(nblocks(n, 256)), dim3(256), 0, s.stream, (const float*)s.partSeg.p,
                     SUM_SEGS, n, out);
  WDG_LAUNCH_CHECK();
}

template <int MODE, int KP>
static void dw(Mpnn& s, const ARows& a, int64_t M, const bf16* dY, float* gW, float* gB) {
  const int64_t rowsPer = (int64_t)DW_CHUNKS * 32;
  const int64_t nblk = std::max<int64_t>((M + rowsPer - 1) / rowsPer, 1);
  hipLaunchKernelGGL((k_mm_dw<MODE, KP>), dim3((unsigned)nblk), dim3(256), 0, s.stream, a, M, dY, s.part.p, s.partB.p);
  WDG_LAUNCH_CHECK();
  sum_partials(s, s.part.p, nblk, (int64_t)KP * H, gW);
  sum_partials(s, s.partB.p, nblk, (int64_t)H, gB);
}

static void forward(Mpnn& s) {
  const int64_t N = s.N, E = s.E;
  for (auto& b : s.h) b.ensure((size_t)std::max<int64_t>(N, 1) * H);
  for (int l = 0; l < 2; l++) {
    s.m[l].ensure((size_t)std::max<int64_t>(E, 1) * H);
    s.agg[l].ensure((size_t)std::max<int64_t>(N, 1) * H);
    s.u[l].ensure((size_t)std::max<int64_t>(N, 1) * H);
  }
  s.out.ensure((size_t)std::max<int64_t>(N, 1) * NOUT);
  s.cflag.ensure((size_t)std::max<int64_t>(N, 1));
  if (N == 0) return;
  hipLaunchKernelGGL(k_contact, dim3(nblocks(N, 256)), dim3(256), 0, s.stream, N, (const int32_t*)s.rowOff.p,
                     (const bf16*)s.X.p, (const bf16*)s.EF.p, s.cflag.p);
  WDG_LAUNCH_CHECK();
  const unsigned gN = fwd_blocks(N), gE = fwd_blocks(E);
  ARows a{};
  a.x = s.X.p;
  a.K = KX;
  hipLaunchKernelGGL((k_mm_fwd<A_PLAIN, false>), dim3(gN), dim3(256), 0, s.stream, a, N,
                     (const bf16*)(s.img.p + s.imgT[0]), (const float*)(s.P.p + s.L.encB), s.h[0].p, (bf16*)nullptr);
  WDG_LAUNCH_CHECK();
  for (int l = 0; l < 2; l++) {
    ARows ae{};
    ae.h = s.h[l].p; ae.ef = s.EF.p; ae.recv = s.recv.p; ae.send = s.send.p; ae.K = KM;
    if (E > 0) {
      hipLaunchKernelGGL((k_mm_fwd<A_EDGE, false>), dim3(gE), dim3(256), 0, s.stream, ae, E,
                         (const bf16*)(s.img.p + s.imgT[1 + 2 * l]), (const float*)(s.P.p + s.L.mB[l]), s.m[l].p,
                         (bf16*)nullptr);
      WDG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_agg, dim3(nblocks(N * 8, 256)), dim3(256), 0, s.stream, N, (const int32_t*)s.rowOff.p,
                       (const bf16*)s.m[l].p, s.agg[l].p);
    WDG_LAUNCH_CHECK();
    ARows au{};
    au.h = s.h[l].p; au.agg = s.agg[l].p; au.K = KU;
    hipLaunchKernelGGL((k_mm_fwd<A_UPD, true>), dim3(gN), dim3(256), 0, s.stream, au, N,
                       (const bf16*)(s.img.p + s.imgT[2 + 2 * l]), (const float*)(s.P.p + s.L.uB[l]), s.u[l].p,
                       s.h[l + 1].p);
    WDG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_dec, dim3(nblocks(N, 256)), dim3(256), 0, s.stream, N, (const bf16*)s.h[2].p,
                     (const float*)(s.P.p + s.L.decW), (const float*)(s.P.p + s.L.decB), s.out.p);
  WDG_LAUNCH_CHECK();
}

// gradients of the loss (mean squared error against s.T) into s.G; returns the loss
static double backward(Mpnn& s) {
  const int64_t N = s.N, E = s.E;
  const int64_t nbN = nblocks(N, 256);
  s.dA.ensure((size_t)N * NOUT);
  s.partLoss.ensure((size_t)nbN);
  s.dhA.ensure((size_t)N * H);
  s.dhB.ensure((size_t)N * H);
  s.dU.ensure((size_t)N * KU);
  s.dZ.ensure((size_t)std::max<int64_t>(E, 1) * KU);
  s.dY.ensure((size_t)N * H);
  s.dYe.ensure((size_t)std::max<int64_t>(E, 1) * H);
  // one partial buffer, sized for the largest user (no reallocation while kernels run)
  {
    const int64_t rowsPer = (int64_t)DW_CHUNKS * 32;
    const int64_t bE = std::max<int64_t>((E + rowsPer - 1) / rowsPer, 1), bN = std::max<int64_t>((N + rowsPer - 1) / rowsPer, 1);
    const size_t need = std::max({(size_t)bE * KM * H, (size_t)bN * KU * H, (size_t)nbN * (H + 1) * NOUT});
    s.part.ensure(need);
    s.partB.ensure((size_t)std::max(bE, bN) * H);
    s.partSeg.ensure((size_t)SUM_SEGS * KM * H);
  }
  WDG_HIP(hipMemsetAsync(s.G.p, 0, s.L.total * sizeof(float), s.stream));
  hipLaunchKernelGGL(k_loss, dim3((unsigned)nbN), dim3(256), 0, s.stream, N, (const float*)s.out.p,
                     (const float*)s.T.p, (const uint8_t*)s.cflag.p, s.dA.p, s.partLoss.p);
  WDG_LAUNCH_CHECK();
  // decoder
  hipLaunchKernelGGL(k_dec_bwd, dim3((unsigned)nbN), dim3(256), 0, s.stream, N, (const bf16*)s.h[2].p,
                     (const float*)(s.P.p + s.L.decW), (const float*)s.dA.p, s.dhA.p, s.part.p);
  WDG_LAUNCH_CHECK();
  sum_partials(s, s.part.p, nbN, (int64_t)(H + 1) * NOUT, s.G.p + s.L.decW);  // decW then decB are adjacent
  float* dh = s.dhA.p;   // gradient of the current layer output
  float* dh2 = s.dhB.p;
  const unsigned gN = (unsigned)((N + 63) / 64), gE = (unsigned)((E + 63) / 64);
  for (int l = 1; l >= 0; l--) {
    // update: u = relu([h, agg] Wu + bu), h_out = h + u
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblocks(N * H, 256)), dim3(256), 0, s.stream, N * H, (const float*)dh,
                       (const bf16*)s.u[l].p, s.dY.p);
    WDG_LAUNCH_CHECK();
    ARows au{};
    au.h = s.h[l].p; au.agg = s.agg[l].p; au.K = KU;
    dw<A_UPD, KU>(s, au, N, s.dY.p, s.G.p + s.L.uW[l], s.G.p + s.L.uB[l]);
    hipLaunchKernelGGL((k_mm_dx<KU>), dim3(gN), dim3(256), 0, s.stream, (const bf16*)s.dY.p, N,
                       (const bf16*)(s.img.p + s.imgR[2 + 2 * l]), s.dU.p);
    WDG_LAUNCH_CHECK();
    // messages
    if (E > 0) {
      hipLaunchKernelGGL(k_edge_dy, dim3(nblocks(E * H, 256)), dim3(256), 0, s.stream, E, (const int32_t*)s.recv.p,
                         (const float*)s.dU.p, (const bf16*)s.m[l].p, s.dYe.p);
      WDG_LAUNCH_CHECK();
      ARows ae{};
      ae.h = s.h[l].p; ae.ef = s.EF.p; ae.recv = s.recv.p; ae.send = s.send.p; ae.K = KM;
      dw<A_EDGE, KM>(s, ae, E, s.dYe.p, s.G.p + s.L.mW[l], s.G.p + s.L.mB[l]);
      // only the node-activation columns of the message input need a gradient
      hipLaunchKernelGGL((k_mm_dx<KU>), dim3(gE), dim3(256), 0, s.stream, (const bf16*)s.dYe.p, E,
                         (const bf16*)(s.img.p + s.imgR[1 + 2 * l]), s.dZ.p);
      WDG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_node_grad, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s.stream, N,
                       (const int32_t*)s.rowOff.p, (const int32_t*)s.rev.p, (const float*)dh, (const float*)s.dU.p,
                       (const float*)s.dZ.p, dh2);
    WDG_LAUNCH_CHECK();
    std::swap(dh, dh2);
  }
  // encoder
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblocks(N * H, 256)), dim3(256), 0, s.stream, N * H, (const float*)dh,
                     (const bf16*)s.h[0].p, s.dY.p);
  WDG_LAUNCH_CHECK();
  ARows ax{};
  ax.x = s.X.p;
  ax.K = KX;
  dw<A_PLAIN, KX>(s, ax, N, s.dY.p, s.G.p + s.L.encW, s.G.p + s.L.encB);
  std::vector<float> pl((size_t)nbN);
  WDG_HIP(hipMemcpyAsync(pl.data(), s.partLoss.p, nbN * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  WDG_HIP(hipStreamSynchronize(s.stream));
  double loss = 0;
  for (float x : pl) loss += x;
  return loss / (double)N;
}

// ------------------------------------------------------------------ graph from an engine context
// Edges between spheres closer than cutoff (ri + rj), from the context's slot list (which must
// hold every such pair: its Verlet distance is at least 0.05 (ri + rj) plus the largest
// displacement since its last rebuild); each sphere's edges in its adjacency order.
// Node n = storage index - nW (spheres only).
struct GraphArgs {
  int64_t nW, nSph;
  const uint32_t *adj, *adjOff, *A, *B, *info;
  const double *px, *py, *pz, *vx, *vy, *vz, *rad;
  double cutoff, floorZ;
  int32_t *deg, *rowOff, *recv, *send, *rev, *slotSide;
  int32_t* edgeOfSlot;  // [2 * nslots]
  bf16 *X, *EF;
};
// a row of K inputs as BF16, in 16-byte stores
template <int K>
__device__ __forceinline__ void store_row(bf16* o, const float* f) {
#pragma unroll
  for (int c = 0; c < K; c += 8) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = (uint32_t)f2bf(f[c + 2 * q]) | ((uint32_t)f2bf(f[c + 2 * q + 1]) << 16);
    *reinterpret_cast<uint4*>(o + c) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}
__device__ __forceinline__ bool edge_keep(const GraphArgs& g, uint32_t s, uint32_t slot, uint32_t& other) {
  if ((g.info[slot] & CI_TYPE_MASK) != CT_SS) return false;
  const uint32_t a = g.A[slot], b = g.B[slot];
  other = a == s ? b : a;
  const double dx = g.px[other] - g.px[s], dy = g.py[other] - g.py[s], dz = g.pz[other] - g.pz[s];
  const double rs = g.rad[s] + g.rad[other];
  return (dx * dx + dy * dy) + dz * dz < (g.cutoff * rs) * (g.cutoff * rs);
}
__global__ void k_mg_count(GraphArgs g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nSph) return;
  const uint32_t s = (uint32_t)(g.nW + t);
  int32_t c = 0;
  uint32_t other;
  for (uint32_t k = g.adjOff[s]; k < g.adjOff[s + 1]; k++) c += edge_keep(g, s, g.adj[k] & ~ADJ_PA, other) ? 1 : 0;
  g.deg[t] = c;
}
// the structure of the graph: each sphere's kept edges in its adjacency order
__global__ void k_mg_fill(GraphArgs g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nSph) return;
  const uint32_t s = (uint32_t)(g.nW + t);
  int32_t e = g.rowOff[t];
  uint32_t other;
  for (uint32_t k = g.adjOff[s]; k < g.adjOff[s + 1]; k++) {
    const uint32_t slot = g.adj[k] & ~ADJ_PA;
    if (!edge_keep(g, s, slot, other)) continue;
    g.send[e] = (int32_t)(other - g.nW);
    g.recv[e] = (int32_t)t;
    const int side = g.A[slot] == s ? 0 : 1;
    g.slotSide[e] = (int32_t)(2 * slot + side);
    g.edgeOfSlot[2 * slot + side] = e;
    e++;
  }
}
// inputs of every edge (threads [0, E)) and node (threads [E, E + nSph)), and the opposite
// edge of each edge; one thread per row, so that the double-precision input arithmetic runs
// on full wavefronts
__global__ __launch_bounds__(256) void k_mg_inputs(GraphArgs g, int64_t E) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < E) {
    const int64_t i = g.nW + g.recv[q], j = g.nW + g.send[q];
    g.rev[q] = g.edgeOfSlot[g.slotSide[q] ^ 1];
    float f[KE];
    for (int c = NEI; c < KE; c++) f[c] = 0.f;
    edge_inputs(g.px[i], g.py[i], g.pz[i], g.px[j], g.py[j], g.pz[j], g.vx[i], g.vy[i], g.vz[i], g.vx[j], g.vy[j],
                g.vz[j], g.rad[i], g.rad[j], f);
    store_row<KE>(g.EF + q * KE, f);
  } else if (q < E + g.nSph) {
    const int64_t t = q - E, s = g.nW + t;
    float x[KX];
    for (int c = NIN; c < KX; c++) x[c] = 0.f;
    node_inputs(g.vx[s], g.vy[s], g.vz[s], g.rad[s], g.pz[s] - g.floorZ, x);
    store_row<KX>(g.X + t * KX, x);
  }
}
// coarse step: v += a dT, x += v dT for the spheres; a is the network's output (in units of
// unit) for a sphere in contact, the field's gravity for one in free flight
__global__ void k_mg_advance(int64_t nW, int64_t nSph, const float* __restrict__ a, const uint8_t* __restrict__ inContact,
                             double dT, double unit, double gx, double gy, double gz, double* px, double* py,
                             double* pz, double* vx, double* vy, double* vz, const uint32_t* flags) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nSph) return;
  const int64_t s = nW + t;
  if (!(flags[s] & F_INTEGRATED)) return;
  const bool c = inContact[t] != 0;
  const double ax = c ? unit * a[t * NOUT] : gx, ay = c ? unit * a[t * NOUT + 1] : gy,
               az = c ? unit * a[t * NOUT + 2] : gz;
  const double v[3] = {vx[s] + ax * dT, vy[s] + ay * dT, vz[s] + az * dT};
  vx[s] = v[0]; vy[s] = v[1]; vz[s] = v[2];
  px[s] = px[s] + v[0] * dT; py[s] = py[s] + v[1] * dT; pz[s] = pz[s] + v[2] * dT;
}
// largest squared displacement since the last rebuild, as ordered bits (non-negative doubles)
__global__ __launch_bounds__(256) void k_mg_disp(int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, const double* __restrict__ lx,
                                                  const double* __restrict__ ly, const double* __restrict__ lz,
                                                  unsigned long long* out) {
  __shared__ double wm[4];
  double m = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double dx = px[i] - lx[i], dy = py[i] - ly[i], dz = pz[i] - lz[i];
    m = fmax(m, (dx * dx + dy * dy) + dz * dz);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  // one device-wide atomic per workgroup
  if (threadIdx.x == 0)
    atomicMax(out, (unsigned long long)__double_as_longlong(fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]))));
}

}  // namespace wdg

using namespace wdg;

struct wdg_mpnn {
  Mpnn s;
};

extern thread_local std::string wdg_mpnn_err;
thread_local std::string wdg_mpnn_err;

template <class F>
static int mguard(F&& f) {
  try {
    f();
    return WDG_OK;
  } catch (const Error& e) {
    wdg_mpnn_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    wdg_mpnn_err = e.what();
    return WDG_E_STATE;
  }
}

// splitmix64 stream for the initial weights
static uint64_t smix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

extern "C" {

const char* wdg_mpnn_last_error(void) { return wdg_mpnn_err.c_str(); }

int wdg_mpnn_create(int device, uint64_t seed, wdg_mpnn** out) {
  if (!out) return WDG_E_ARG;
  *out = nullptr;
  auto* m = new wdg_mpnn();
  int rc = mguard([&] {
    Mpnn& s = m->s;
    s.device = device;
    WDG_HIP(hipSetDevice(device));
    WDG_HIP(hipStreamCreate(&s.stream));
    WDG_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    const size_t n = s.L.total;
    s.P.ensure(n); s.G.ensure(n); s.Am.ensure(n); s.Av.ensure(n);
    // Glorot-uniform weights over the used input rows, zero biases and padding rows (the
    // decoder below)
    std::vector<float> p(n, 0.f);
    uint64_t st = seed;
    auto fill = [&](size_t off, int K, int Kused, int O) {
      const float lim = std::sqrt(6.f / (float)(Kused + O));
      for (int k = 0; k < K; k++)
        for (int o = 0; o < O; o++) {
          const double u = (double)(smix(st) >> 11) * (1.0 / 9007199254740992.0);
          p[off + (size_t)k * O + o] = k < Kused ? (float)((2 * u - 1) * lim) : 0.f;
        }
    };
    fill(s.L.encW, KX, NIN, H);
    for (int l = 0; l < 2; l++) {
      fill(s.L.mW[l], KM, 2 * H + NEI, H);
      // rows of the unused edge-input columns are zero
      for (int k = 2 * H + NEI; k < KM; k++)
        for (int o = 0; o < H; o++) p[s.L.mW[l] + (size_t)k * H + o] = 0.f;
      fill(s.L.uW[l], KU, KU, H);
    }
    // the decoder starts at zero weights and a bias of free fall: the untrained network is the
    // ballistic coarse step, and training learns the contact accelerations on top of it
    p[s.L.decB + 2] = (float)(-9.81 / MPNN_A0);
    WDG_HIP(hipMemcpy(s.P.p, p.data(), n * 4, hipMemcpyHostToDevice));
    WDG_HIP(hipMemset(s.Am.p, 0, n * 4));
    WDG_HIP(hipMemset(s.Av.p, 0, n * 4));
    WDG_HIP(hipMemset(s.G.p, 0, n * 4));
    images(s);
    WDG_HIP(hipStreamSynchronize(s.stream));
  });
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return WDG_OK;
}

int wdg_mpnn_destroy(wdg_mpnn* m) {
  if (!m) return WDG_OK;
  Mpnn& s = m->s;
  DBuf<float>* fb[] = {&s.P, &s.G, &s.Am, &s.Av, &s.T, &s.out, &s.dA, &s.dhA, &s.dhB, &s.dU, &s.dZ, &s.part, &s.partB,
                       &s.partLoss, &s.partSeg};
  for (auto* b : fb) b->release();
  DBuf<bf16>* bb[] = {&s.img, &s.X, &s.EF, &s.h[0], &s.h[1], &s.h[2], &s.m[0], &s.m[1], &s.agg[0], &s.agg[1],
                      &s.u[0], &s.u[1], &s.dY, &s.dYe};
  for (auto* b : bb) b->release();
  s.cflag.release();
  s.rowOff.release(); s.recv.release(); s.send.release(); s.rev.release();
  s.deg.release(); s.slotSide.release(); s.edgeOfSlot.release(); s.degScan.release(); s.nodeUid.release();
  for (auto& k : s.kept) k.release();
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  delete m;
  return WDG_OK;
}

int wdg_mpnn_num_params(wdg_mpnn* m, int64_t* n) {
  if (!m || !n) return WDG_E_ARG;
  *n = (int64_t)m->s.L.total;
  return WDG_OK;
}

int wdg_mpnn_get_params(wdg_mpnn* m, float* p, float* g) {
  if (!m) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    WDG_HIP(hipStreamSynchronize(s.stream));
    if (p) WDG_HIP(hipMemcpy(p, s.P.p, s.L.total * 4, hipMemcpyDeviceToHost));
    if (g) WDG_HIP(hipMemcpy(g, s.G.p, s.L.total * 4, hipMemcpyDeviceToHost));
  });
}

int wdg_mpnn_set_params(wdg_mpnn* m, const float* p) {
  if (!m || !p) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    WDG_HIP(hipMemcpy(s.P.p, p, s.L.total * 4, hipMemcpyHostToDevice));
    images(s);
    WDG_HIP(hipStreamSynchronize(s.stream));
  });
}

/* graph from host arrays: x [N x 7] node inputs, CSR by receiver rowOff[N+1], send[E], rev[E],
 * ef [E x 15] edge inputs, targets t [N x 3] (may be NULL) */
int wdg_mpnn_set_graph(wdg_mpnn* m, int64_t N, int64_t E, const float* x, const int32_t* rowOff,
                       const int32_t* send, const int32_t* rev, const float* ef, const float* t) {
  if (!m || N < 0 || E < 0 || (N > 0 && (!x || !rowOff)) || (E > 0 && (!send || !rev || !ef))) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    if (rowOff[0] != 0 || rowOff[N] != E) throw Error(WDG_E_ARG, "rowOff must run from 0 to E");
    std::vector<bf16> X((size_t)std::max<int64_t>(N, 1) * KX, 0), F((size_t)std::max<int64_t>(E, 1) * KE, 0);
    std::vector<int32_t> recv((size_t)std::max<int64_t>(E, 1));
    for (int64_t i = 0; i < N; i++) {
      if (rowOff[i + 1] < rowOff[i]) throw Error(WDG_E_ARG, "rowOff must not decrease");
      for (int c = 0; c < NIN; c++) X[i * KX + c] = f2bf(x[i * NIN + c]);
      for (int32_t e = rowOff[i]; e < rowOff[i + 1]; e++) recv[e] = (int32_t)i;
    }
    for (int64_t e = 0; e < E; e++) {
      if (send[e] < 0 || send[e] >= N || rev[e] < 0 || rev[e] >= E) throw Error(WDG_E_ARG, "edge index out of range");
      for (int c = 0; c < NEI; c++) F[e * KE + c] = f2bf(ef[e * NEI + c]);
    }
    s.N = N;
    s.E = E;
    s.outRow.clear();
    s.fromCtx = nullptr;
    s.X.ensure(X.size()); s.EF.ensure(F.size());
    s.rowOff.ensure((size_t)N + 1); s.recv.ensure(recv.size()); s.send.ensure(recv.size()); s.rev.ensure(recv.size());
    s.T.ensure((size_t)std::max<int64_t>(N, 1) * NOUT);
    WDG_HIP(hipMemcpy(s.X.p, X.data(), X.size() * 2, hipMemcpyHostToDevice));
    WDG_HIP(hipMemcpy(s.EF.p, F.data(), F.size() * 2, hipMemcpyHostToDevice));
    WDG_HIP(hipMemcpy(s.rowOff.p, rowOff, ((size_t)N + 1) * 4, hipMemcpyHostToDevice));
    if (E > 0) {
      WDG_HIP(hipMemcpy(s.recv.p, recv.data(), E * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(s.send.p, send, E * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(s.rev.p, rev, E * 4, hipMemcpyHostToDevice));
    }
    if (t) WDG_HIP(hipMemcpy(s.T.p, t, (size_t)N * NOUT * 4, hipMemcpyHostToDevice));
  });
}

/* forward pass; out [N x 3] (may be NULL) */
int wdg_mpnn_forward(wdg_mpnn* m, float* out) {
  if (!m) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    forward(s);
    if (out) out_rows(s);
    if (out && s.N > 0) {
      std::vector<float> o((size_t)s.N * NOUT);
      WDG_HIP(hipMemcpyAsync(o.data(), s.out.p, o.size() * 4, hipMemcpyDeviceToHost, s.stream));
      WDG_HIP(hipStreamSynchronize(s.stream));
      for (int64_t t = 0; t < s.N; t++) {
        const int64_t r = s.outRow.empty() ? t : s.outRow[t];
        for (int c2 = 0; c2 < NOUT; c2++) out[r * NOUT + c2] = o[t * NOUT + c2];
      }
    }
    // without an output the pass stays queued: wdg_mpnn_advance orders the engine's stream after it
    if (!s.fromCtx) WDG_HIP(hipStreamSynchronize(s.stream));
  });
}

/* targets t [N x 3] of the current graph, in its output order */
int wdg_mpnn_set_targets(wdg_mpnn* m, const float* t) {
  if (!m || !t) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    out_rows(s);
    std::vector<float> T((size_t)std::max<int64_t>(s.N, 1) * NOUT);
    for (int64_t q = 0; q < s.N; q++) {
      const int64_t r = s.outRow.empty() ? q : s.outRow[q];
      for (int c2 = 0; c2 < NOUT; c2++) T[q * NOUT + c2] = t[r * NOUT + c2];
    }
    s.T.ensure(T.size());
    WDG_HIP(hipMemcpyAsync(s.T.p, T.data(), (size_t)s.N * NOUT * 4, hipMemcpyHostToDevice, s.stream));
    WDG_HIP(hipStreamSynchronize(s.stream));
  });
}

/* the current graph: sizes, and (each array may be NULL) its CSR, inputs as stored (BF16
 * widened to float) and the output row of each node */
int wdg_mpnn_get_graph(wdg_mpnn* m, int64_t* N, int64_t* E, int32_t* rowOff, int32_t* send, int32_t* rev, float* x,
                       float* ef, int64_t* nodeRow) {
  if (!m) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    WDG_HIP(hipStreamSynchronize(s.stream));
    if (N) *N = s.N;
    if (E) *E = s.E;
    if (rowOff) WDG_HIP(hipMemcpy(rowOff, s.rowOff.p, ((size_t)s.N + 1) * 4, hipMemcpyDeviceToHost));
    if (send && s.E) WDG_HIP(hipMemcpy(send, s.send.p, (size_t)s.E * 4, hipMemcpyDeviceToHost));
    if (rev && s.E) WDG_HIP(hipMemcpy(rev, s.rev.p, (size_t)s.E * 4, hipMemcpyDeviceToHost));
    auto widen = [&](const DBuf<bf16>& b, int64_t rows, int K, int used, float* out) {
      std::vector<bf16> h((size_t)rows * K);
      if (rows) WDG_HIP(hipMemcpy(h.data(), b.p, h.size() * 2, hipMemcpyDeviceToHost));
      for (int64_t r = 0; r < rows; r++)
        for (int c2 = 0; c2 < used; c2++) out[r * used + c2] = bf2f(h[(size_t)r * K + c2]);
    };
    if (x) widen(s.X, s.N, KX, NIN, x);
    if (ef) widen(s.EF, s.E, KE, NEI, ef);
    if (nodeRow) {
      out_rows(s);
      for (int64_t q = 0; q < s.N; q++) nodeRow[q] = s.outRow.empty() ? q : s.outRow[q];
    }
  });
}

/* keep the current graph (inputs, structure, targets) on the device under slot */
int wdg_mpnn_keep_graph(wdg_mpnn* m, int32_t slot) {
  if (!m || slot < 0) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    if ((size_t)slot >= s.kept.size()) s.kept.resize((size_t)slot + 1);
    Mpnn::Kept& k = s.kept[slot];
    copy_buf(k.rowOff, s.rowOff, (size_t)s.N + 1, s.stream);
    copy_buf(k.recv, s.recv, (size_t)s.E, s.stream);
    copy_buf(k.send, s.send, (size_t)s.E, s.stream);
    copy_buf(k.rev, s.rev, (size_t)s.E, s.stream);
    copy_buf(k.X, s.X, (size_t)s.N * KX, s.stream);
    copy_buf(k.EF, s.EF, (size_t)s.E * KE, s.stream);
    copy_buf(k.T, s.T, (size_t)s.N * NOUT, s.stream);
    k.N = s.N;
    k.E = s.E;
    WDG_HIP(hipStreamSynchronize(s.stream));
  });
}

/* make a kept graph current (outputs in its node order) */
int wdg_mpnn_use_graph(wdg_mpnn* m, int32_t slot) {
  if (!m || slot < 0) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    if ((size_t)slot >= s.kept.size() || s.kept[slot].N < 0) throw Error(WDG_E_ARG, "no graph kept in this slot");
    Mpnn::Kept& k = s.kept[slot];
    s.N = k.N;
    s.E = k.E;
    copy_buf(s.rowOff, k.rowOff, (size_t)s.N + 1, s.stream);
    copy_buf(s.recv, k.recv, (size_t)s.E, s.stream);
    copy_buf(s.send, k.send, (size_t)s.E, s.stream);
    copy_buf(s.rev, k.rev, (size_t)s.E, s.stream);
    copy_buf(s.X, k.X, (size_t)s.N * KX, s.stream);
    copy_buf(s.EF, k.EF, (size_t)s.E * KE, s.stream);
    copy_buf(s.T, k.T, (size_t)s.N * NOUT, s.stream);
    s.outRow.clear();
    s.fromCtx = nullptr;
  });
}

/* forward + backward against the targets of the graph: loss, gradients (wdg_mpnn_get_params) */
int wdg_mpnn_grad(wdg_mpnn* m, double* loss) {
  if (!m) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    if (s.N == 0) throw Error(WDG_E_STATE, "empty graph");
    forward(s);
    const double l = backward(s);
    if (loss) *loss = l;
  });
}

/* one Adam step (lr, beta1, beta2, eps) with the current gradients */
int wdg_mpnn_adam(wdg_mpnn* m, double lr, double b1, double b2, double eps) {
  if (!m) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    s.adamStep++;
    const double c1 = 1 - std::pow(b1, (double)s.adamStep), c2 = 1 - std::pow(b2, (double)s.adamStep);
    hipLaunchKernelGGL(k_adam, dim3(nblocks((int64_t)s.L.total, 256)), dim3(256), 0, s.stream, (int64_t)s.L.total,
                       s.P.p, (const float*)s.G.p, s.Am.p, s.Av.p, (float)lr, (float)b1, (float)b2, (float)eps,
                       (float)c1, (float)c2);
    WDG_LAUNCH_CHECK();
    images(s);
    WDG_HIP(hipStreamSynchronize(s.stream));
  });
}

int wdg_mpnn_graph_from_ctx(wdg_mpnn* m, wdg_ctx* ctx, double cutoff, double floorZ, double maxDisp, int force,
                            int64_t* nNodes, int64_t* nEdges, int* rebuilt) {
  if (!m || !ctx || !(cutoff > 0)) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    Ctx& c = ctx_of(ctx);
    if (c.field.periodic) throw Error(WDG_E_ARG, "the learned step supports aperiodic scenes");
    if (c.forcesPending) flush_forces(c);
    const int64_t nSph = c.n - c.nW;
    // the network's kernels of the previous graph (on its own stream) end before the graph changes
    WDG_HIP(hipEventRecord(s.done, s.stream));
    WDG_HIP(hipStreamWaitEvent(c.stream, s.done, 0));
    s.deg.ensure((size_t)nSph + 1);
    s.rowOff.ensure((size_t)nSph + 1);
    GraphArgs g{};
    g.nW = c.nW; g.nSph = nSph;
    g.cutoff = cutoff; g.floorZ = floorZ;
    g.deg = s.deg.p; g.rowOff = s.rowOff.p;
    // the context's arrays as they are now: a refresh permutes the particles into buffers of its
    // own, so these are bound again after every collider call
    auto bind = [&]() {
      Contacts& C = c.con();
      g.adj = C.adj.p; g.adjOff = C.adjOff.p; g.A = C.A.p; g.B = C.B.p; g.info = C.info.p;
      g.px = c.pos[0].p; g.py = c.pos[1].p; g.pz = c.pos[2].p;
      g.vx = c.vel[0].p; g.vy = c.vel[1].p; g.vz = c.vel[2].p; g.rad = c.radius.p;
    };
    // edge counts over the current slot list, and in the same round trip the largest
    // displacement since its refresh (the slot list holds every edge while that stays below maxDisp)
    auto count = [&](bool disp, double& d2) -> int64_t {
      bind();
      unsigned long long* d = reinterpret_cast<unsigned long long*>(c.red.p);
      if (disp) {
        WDG_HIP(hipMemsetAsync(d, 0, 8, c.stream));
        hipLaunchKernelGGL(k_mg_disp, dim3((unsigned)std::min<int64_t>(nblocks(nSph, 256), 256)), dim3(256), 0,
                           c.stream, nSph, (const double*)(c.pos[0].p + c.nW), (const double*)(c.pos[1].p + c.nW),
                           (const double*)(c.pos[2].p + c.nW), (const double*)(c.lastPos[0].p + c.nW),
                           (const double*)(c.lastPos[1].p + c.nW), (const double*)(c.lastPos[2].p + c.nW), d);
        WDG_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_mg_count, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, g);
      WDG_LAUNCH_CHECK();
      scan_exclusive(c, reinterpret_cast<const uint32_t*>(s.deg.p), reinterpret_cast<uint32_t*>(s.rowOff.p), nSph,
                     reinterpret_cast<uint32_t*>(s.rowOff.p + nSph));
      uint32_t hE = 0;
      unsigned long long hb = 0;
      WDG_HIP(hipMemcpyAsync(&hE, s.rowOff.p + nSph, 4, hipMemcpyDeviceToHost, c.stream));
      if (disp) WDG_HIP(hipMemcpyAsync(&hb, d, 8, hipMemcpyDeviceToHost, c.stream));
      WDG_HIP(hipStreamSynchronize(c.stream));
      memcpy(&d2, &hb, 8);
      return (int64_t)hE;
    };
    if (std::isnan(c.verletDist)) throw Error(WDG_E_STATE, "run the collider once before building graphs");
    // the slot list holds every pair closer than r1 + r2 + 2 verletDist at its refresh, so it
    // holds every edge while no sphere has moved farther than verletDist - need since then
    const double need = 0.5 * (cutoff - 1.) * 2. * c.rmax;
    const double allow = c.verletDist - need;
    bool rb = force != 0 || !c.boundsValid || allow < 0;
    int64_t E = 0;
    if (!rb && nSph > 0) {
      double d2 = 0.;
      E = count(true, d2);
      rb = d2 > allow * allow;
    }
    if (rb) {
      // rebuild with room for maxDisp of motion before the next one
      c.verletDist = std::max(c.verletDist, need + std::max(maxDisp, 0.));
      wdg_collide_params p{c.verletDist, 1};
      collide(c, &p);
      double d2;
      if (nSph > 0) E = count(false, d2);
    }
    if (rebuilt) *rebuilt = rb ? 1 : 0;
    bind();
    Contacts& C = c.con();
    s.edgeOfSlot.ensure((size_t)std::max<int64_t>(2 * C.n, 2));
    s.recv.ensure((size_t)std::max<int64_t>(E, 1)); s.send.ensure((size_t)std::max<int64_t>(E, 1));
    s.rev.ensure((size_t)std::max<int64_t>(E, 1)); s.slotSide.ensure((size_t)std::max<int64_t>(E, 1));
    s.X.ensure((size_t)std::max<int64_t>(nSph, 1) * KX);
    s.EF.ensure((size_t)std::max<int64_t>(E, 1) * KE);
    s.T.ensure((size_t)std::max<int64_t>(nSph, 1) * NOUT);
    g.recv = s.recv.p; g.send = s.send.p; g.rev = s.rev.p; g.slotSide = s.slotSide.p; g.edgeOfSlot = s.edgeOfSlot.p;
    g.X = s.X.p; g.EF = s.EF.p;
    if (nSph > 0) {
      hipLaunchKernelGGL(k_mg_fill, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, g);
      WDG_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mg_inputs, dim3(nblocks(E + nSph, 256)), dim3(256), 0, c.stream, g, E);
      WDG_LAUNCH_CHECK();
    }
    // the user ids of the nodes change only with a refresh (or an upload) of the context
    if (s.uidCtx != c.serial || s.uidRefresh != c.nRefresh || s.uidUploads != c.uploads || s.uidN != nSph) {
      s.nodeUid.ensure((size_t)std::max<int64_t>(nSph, 1));
      if (nSph > 0)
        WDG_HIP(hipMemcpyAsync(s.nodeUid.p, c.uid.p + c.nW, nSph * 4, hipMemcpyDeviceToDevice, c.stream));
      s.uidCtx = c.serial;
      s.uidRefresh = c.nRefresh;
      s.uidUploads = c.uploads;
      s.uidN = nSph;
    }
    // the network's next kernels (its own stream) start after the graph is built
    WDG_HIP(hipEventRecord(s.done, c.stream));
    WDG_HIP(hipStreamWaitEvent(s.stream, s.done, 0));
    // output rows (spheres in user-id order) are found when an output is asked for
    s.outRow.clear();
    s.fromCtx = &c;
    s.N = nSph;
    s.E = E;
    if (nNodes) *nNodes = nSph;
    if (nEdges) *nEdges = E;
  });
}

int wdg_mpnn_advance(wdg_mpnn* m, wdg_ctx* ctx, double dT) {
  if (!m || !ctx) return WDG_E_ARG;
  return mguard([&] {
    Mpnn& s = m->s;
    Ctx& c = ctx_of(ctx);
    const int64_t nSph = c.n - c.nW;
    if (s.N != nSph) throw Error(WDG_E_STATE, "the network's graph is not of this context");
    if (s.fromCtx != &c) throw Error(WDG_E_STATE, "the network's graph was not built from this context");
    // the forward pass (network stream) ends before the engine's stream moves the spheres
    WDG_HIP(hipEventRecord(s.done, s.stream));
    WDG_HIP(hipStreamWaitEvent(c.stream, s.done, 0));
    if (c.forcesPending) flush_forces(c);
    // the contact history describes the state before the coarse step, which the network
    // replaces: the next contact loop starts it afresh
    clear_history(c);
    if (nSph > 0) {
      hipLaunchKernelGGL(k_mg_advance, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, c.nW, nSph,
                         (const float*)s.out.p, (const uint8_t*)s.cflag.p, dT, MPNN_A0, c.field.gravity[0],
                         c.field.gravity[1], c.field.gravity[2], c.pos[0].p, c.pos[1].p, c.pos[2].p, c.vel[0].p, c.vel[1].p,
                         c.vel[2].p, (const uint32_t*)c.flags.p);
      WDG_LAUNCH_CHECK();
    }
    c.dirtyKnown = false;
  });
}

}  // extern "C"
