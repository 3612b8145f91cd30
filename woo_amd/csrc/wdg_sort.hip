// Stable LSD radix sort of (key, value) pairs and exclusive prefix scan, for gfx950.
//
// Radix sort: 8-bit digits, one pass = histogram kernel (LDS bins per tile),
// exclusive scan of the digit-major histogram, stable scatter. The scatter
// ranks equal digits inside a 64-lane wavefront with eight ballots (lanes whose
// digit matches in every bit), then across the four wavefronts of the block
// through LDS counters, so the output keeps input order among equal keys.
#include "wdg_internal.hpp"

namespace wdg {

#if defined(__HIP_DEVICE_COMPILE__) && defined(__AMDGCN_WAVEFRONT_SIZE) && __AMDGCN_WAVEFRONT_SIZE != 64
#error "the ballot-based ranking assumes 64-lane wavefronts"
#endif

constexpr int SCAN_BS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BS * SCAN_ITEMS;

// ------------------------------------------------------------------ scan
__global__ __launch_bounds__(SCAN_BS) void k_scan_block(const uint32_t* __restrict__ in, uint32_t* out, int64_t n,
                                                         uint32_t* blockSums) {
  __shared__ uint32_t s[SCAN_BS];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t v[SCAN_ITEMS];
  // striped load (coalesced), then each thread owns SCAN_ITEMS consecutive elements via LDS transpose
  __shared__ uint32_t tile[SCAN_TILE];
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t idx = base + (int64_t)i * SCAN_BS + tid;
    tile[i * SCAN_BS + tid] = idx < n ? in[idx] : 0u;
  }
  __syncthreads();
  uint32_t sum = 0;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    v[i] = sum;
    sum += tile[tid * SCAN_ITEMS + i];
  }
  s[tid] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 256 thread sums
  for (int off = 1; off < SCAN_BS; off <<= 1) {
    uint32_t t = tid >= off ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  uint32_t excl = s[tid] - sum;
  for (int i = 0; i < SCAN_ITEMS; i++) tile[tid * SCAN_ITEMS + i] = v[i] + excl;
  __syncthreads();
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t idx = base + (int64_t)i * SCAN_BS + tid;
    if (idx < n) out[idx] = tile[i * SCAN_BS + tid];
  }
  if (tid == SCAN_BS - 1) blockSums[blockIdx.x] = s[SCAN_BS - 1];
}

__global__ void k_scan_add(uint32_t* out, int64_t n, const uint32_t* __restrict__ blockOff) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += blockOff[idx / SCAN_TILE];
}

// single-workgroup scan for short arrays (one launch); in may equal out
constexpr int S1_BS = 1024;
constexpr int S1_ITEMS = 8;
constexpr int64_t S1_MAX = 8192;
__global__ __launch_bounds__(S1_BS) void k_scan_one(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* dTotal) {
  __shared__ uint32_t wsum[S1_BS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t carry = 0;
  for (int64_t base = 0; base < n; base += S1_BS * S1_ITEMS) {
    uint32_t v[S1_ITEMS], sum = 0;
    for (int i = 0; i < S1_ITEMS; i++) {
      const int64_t idx = base + (int64_t)tid * S1_ITEMS + i;
      const uint32_t x = idx < n ? in[idx] : 0u;
      v[i] = sum;
      sum += x;
    }
    uint32_t inc = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(inc, off);
      if (lane >= off) inc += t;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (int w = 0; w < S1_BS / 64; w++) {
      const uint32_t s = wsum[w];
      if (w < wv) wpre += s;
      tot += s;
    }
    const uint32_t excl = carry + wpre + (inc - sum);
    for (int i = 0; i < S1_ITEMS; i++) {
      const int64_t idx = base + (int64_t)tid * S1_ITEMS + i;
      if (idx < n) out[idx] = v[i] + excl;
    }
    carry += tot;
    __syncthreads();
  }
  if (tid == 0) *dTotal = carry;
}

static void scan_rec(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* aux, uint32_t* dTotal) {
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SCAN_BS), 0, c.stream, in, out, n, dTotal);
    WDG_LAUNCH_CHECK();
    return;
  }
  uint32_t* sums = aux;
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SCAN_BS), 0, c.stream, in, out, n, sums);
  WDG_LAUNCH_CHECK();
  scan_rec(c, sums, sums, nb, aux + nb + 64, dTotal);
  hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, out, n, (const uint32_t*)sums);
  WDG_LAUNCH_CHECK();
}

void scan_exclusive(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* dTotal) {
  if (n <= 0) {
    WDG_HIP(hipMemsetAsync(dTotal, 0, sizeof(uint32_t), c.stream));
    return;
  }
  if (n <= S1_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(S1_BS), 0, c.stream, in, out, n, dTotal);
    WDG_LAUNCH_CHECK();
    return;
  }
  int64_t need = 0, m = n;
  while (m > 1) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    need += m + 64;
  }
  c.scanAux.ensure((size_t)need + 64);
  scan_rec(c, in, out, n, c.scanAux.p, dTotal);
}

uint32_t read_u32(Ctx& c, const uint32_t* d) {
  WDG_HIP(hipMemcpyAsync(c.hflags + 8, d, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  return c.hflags[8];
}

// ------------------------------------------------------------------ radix sort
constexpr int RS_BS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BS * RS_ITEMS;
constexpr int RS_WAVES = RS_BS / 64;

template <class K>
__global__ __launch_bounds__(RS_BS) void k_rs_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                    uint32_t* hist, int nblk) {
  __shared__ uint32_t h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int i = 0; i < RS_ITEMS; i++) {
    int64_t idx = base + (int64_t)i * RS_BS + tid;
    if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)tid * nblk + blockIdx.x] = h[tid];
}

template <class K>
__global__ __launch_bounds__(RS_BS) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       K* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                       int shift, const uint32_t* __restrict__ histScan, int nblk) {
  __shared__ uint32_t base[256], running[256], wcount[RS_WAVES][256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  base[tid] = histScan[(int64_t)tid * nblk + blockIdx.x];
  running[tid] = 0;
  for (int w = 0; w < RS_WAVES; w++) wcount[w][tid] = 0;
  __syncthreads();
  const int64_t tbase = (int64_t)blockIdx.x * RS_TILE;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r = 0; r < RS_ITEMS; r++) {
    int64_t idx = tbase + (int64_t)r * RS_BS + tid;
    bool valid = idx < n;
    K key = valid ? kin[idx] : K(0);
    uint32_t val = valid ? vin[idx] : 0u;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t rank = 0;
    if (valid) {
      rank = (uint32_t)__popcll(peers & lt);
      int leader = 63 - __clzll((long long)peers);
      if (lane == leader) wcount[wave][d] = (uint32_t)__popcll(peers);
    }
    __syncthreads();
    if (valid) {
      uint32_t off = running[d] + rank;
      for (int w = 0; w < wave; w++) off += wcount[w][d];
      uint32_t dst = base[d] + off;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < RS_WAVES; w++) {
      add += wcount[w][tid];
      wcount[w][tid] = 0;
    }
    running[tid] += add;
    __syncthreads();
  }
}

template <class K>
void radix_sort_pairs(Ctx& c, K* keys, uint32_t* vals, K* keysTmp, uint32_t* valsTmp, int64_t n, int bits) {
  if (n <= 1) return;
  int passes = (bits + 7) / 8;
  int nblk = (int)((n + RS_TILE - 1) / RS_TILE);
  c.sortHist.ensure((size_t)256 * nblk + 64);
  K* ka = keys;
  K* kb = keysTmp;
  uint32_t* va = vals;
  uint32_t* vb = valsTmp;
  for (int p = 0; p < passes; p++) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist<K>, dim3(nblk), dim3(RS_BS), 0, c.stream, (const K*)ka, n, shift, c.sortHist.p, nblk);
    WDG_LAUNCH_CHECK();
    scan_exclusive(c, c.sortHist.p, c.sortHist.p, (int64_t)256 * nblk, c.dflags.p + 3);
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nblk), dim3(RS_BS), 0, c.stream, (const K*)ka, (const uint32_t*)va, kb, vb,
                       n, shift, (const uint32_t*)c.sortHist.p, nblk);
    WDG_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    WDG_HIP(hipMemcpyAsync(keys, ka, n * sizeof(K), hipMemcpyDeviceToDevice, c.stream));
    WDG_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c.stream));
  }
}

template void radix_sort_pairs<uint32_t>(Ctx&, uint32_t*, uint32_t*, uint32_t*, uint32_t*, int64_t, int);
template void radix_sort_pairs<uint64_t>(Ctx&, uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t, int);

}  // namespace wdg
