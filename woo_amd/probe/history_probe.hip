// Measurement probe, not part of the engine: the cost of reaching each contact's history
// through a device hash keyed by the particle ids, against the engine's layout (state packed
// densely in slot order and read as a stream). Both kernels read the 11 history doubles of
// every contact (previous normal 3, local y axis 3, contact point 3, shear force 2) and write
// 10 back, with a trivial update; only the way the state is found differs. bench.py
// --history-probe runs both on the real contacts of the benchmark's fixed state.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace {

constexpr int NS = 11;            // history doubles per contact
constexpr int REC = 12;           // doubles per hash record (11 + pad: 96 B)
constexpr uint64_t EMPTY = ~0ull;

__device__ __forceinline__ uint32_t mix(uint64_t x, uint32_t mask) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return (uint32_t)x & mask;
}

// the engine's way: contact i's state at index i of NS arrays
__global__ __launch_bounds__(256) void k_hist_stream(int64_t n, double* __restrict__ st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) v[q] = st[q * n + i];
  double s = 0.;
#pragma unroll
  for (int q = 0; q < NS; q++) s += v[q];
#pragma unroll
  for (int q = 0; q < NS - 1; q++) st[q * n + i] = v[q] + 1e-12 * s;
}

// the hash: linear probing over the keys (8 B each, 16 to a cache line), then the record (96 B)
__global__ __launch_bounds__(256) void k_hist_hash(int64_t n, const uint64_t* __restrict__ want,
                                                   const uint64_t* __restrict__ keys, uint32_t mask,
                                                   double* __restrict__ rec, uint32_t* miss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = want[i];
  uint32_t h = mix(k, mask);
  uint32_t probes = 0;
  while (keys[h] != k) {
    if (keys[h] == EMPTY || ++probes > mask) {
      atomicAdd(miss, 1u);
      return;
    }
    h = (h + 1) & mask;
  }
  double* r = rec + (int64_t)REC * h;
  double v[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) v[q] = r[q];
  double s = 0.;
#pragma unroll
  for (int q = 0; q < NS; q++) s += v[q];
#pragma unroll
  for (int q = 0; q < NS - 1; q++) r[q] = v[q] + 1e-12 * s;
}

__global__ void k_hash_insert(int64_t n, const uint64_t* __restrict__ want, unsigned long long* keys, uint32_t mask,
                              uint32_t* fail) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = want[i];
  uint32_t h = mix(k, mask);
  for (uint32_t p = 0; p <= mask; p++) {
    const unsigned long long prev = atomicCAS(&keys[h], (unsigned long long)EMPTY, k);
    if (prev == (unsigned long long)EMPTY || prev == k) return;
    h = (h + 1) & mask;
  }
  atomicAdd(fail, 1u);
}

#define HP_CHECK(x)                 \
  do {                              \
    if ((x) != hipSuccess) return -1; \
  } while (0)

}  // namespace

extern "C" {

// n contact keys (id_a << 32 | id_b); mode 0: the engine's streamed state, 1: the hash (load
// factor 1/2); reps timed launches after one untimed; *ms = mean time per launch, *miss = keys
// the hash did not find (0 expected). Returns 0, or -1 on a HIP error.
int hp_run(const uint64_t* hkeys, int64_t n, int mode, int reps, double* ms, uint32_t* miss) {
  if (n <= 0 || reps <= 0) return -1;
  uint64_t* want = nullptr;
  uint32_t* cnt = nullptr;
  HP_CHECK(hipMalloc(&want, n * sizeof(uint64_t)));
  HP_CHECK(hipMalloc(&cnt, 2 * sizeof(uint32_t)));
  HP_CHECK(hipMemcpy(want, hkeys, n * sizeof(uint64_t), hipMemcpyHostToDevice));
  HP_CHECK(hipMemset(cnt, 0, 2 * sizeof(uint32_t)));
  double* st = nullptr;
  unsigned long long* keys = nullptr;
  uint32_t mask = 0;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (mode == 0) {
    HP_CHECK(hipMalloc(&st, (size_t)NS * n * sizeof(double)));
    HP_CHECK(hipMemset(st, 0, (size_t)NS * n * sizeof(double)));
  } else {
    uint64_t cap = 1024;
    while (cap < 2 * (uint64_t)n) cap <<= 1;
    mask = (uint32_t)(cap - 1);
    HP_CHECK(hipMalloc(&keys, cap * sizeof(uint64_t)));
    HP_CHECK(hipMemset(keys, 0xFF, cap * sizeof(uint64_t)));
    HP_CHECK(hipMalloc(&st, cap * REC * sizeof(double)));
    HP_CHECK(hipMemset(st, 0, cap * REC * sizeof(double)));
    hipLaunchKernelGGL(k_hash_insert, grid, block, 0, 0, n, (const uint64_t*)want, keys, mask, cnt + 1);
    HP_CHECK(hipGetLastError());
  }
  hipEvent_t e0, e1;
  HP_CHECK(hipEventCreate(&e0));
  HP_CHECK(hipEventCreate(&e1));
  float total = 0.f;
  for (int r = 0; r <= reps; r++) {
    HP_CHECK(hipEventRecord(e0, 0));
    if (mode == 0) hipLaunchKernelGGL(k_hist_stream, grid, block, 0, 0, n, st);
    else hipLaunchKernelGGL(k_hist_hash, grid, block, 0, 0, n, (const uint64_t*)want, (const uint64_t*)keys, mask, st,
                            cnt);
    HP_CHECK(hipGetLastError());
    HP_CHECK(hipEventRecord(e1, 0));
    HP_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HP_CHECK(hipEventElapsedTime(&t, e0, e1));
    if (r > 0) total += t;
  }
  uint32_t hc[2] = {0, 0};
  HP_CHECK(hipMemcpy(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost));
  *ms = total / reps;
  *miss = hc[0] + hc[1];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(want);
  (void)hipFree(cnt);
  (void)hipFree(st);
  if (keys) (void)hipFree(keys);
  return 0;
}

}  // extern "C"
