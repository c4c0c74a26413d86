// Records what gfx950's stochastic-rounding conversions return: for each
// (x, r) pair, v_cvt_sr_bf16_f32 and v_cvt_sr_f16_f32 (result in the low
// half).  The pairs are every combination of special FP32 patterns (zeros,
// FP32 and FP16 subnormals, the FP16 and BF16 overflow edges, inf, nan
// payloads, values whose discarded bits sit at the rounding edges) with
// special random operands, plus 2^20 seeded random pairs.  Run on an MI355X by
// tests/golden/gen_sr_hw.py, which keeps a subset as sr_hw_vectors.npz: the
// oracle's stochastic rounding (oracle_sr_round) is checked bitwise against
// it on the CPU.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <vector>
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 half2_ __attribute__((ext_vector_type(2)));

__global__ void probe(const uint32_t *x, const uint32_t *r, uint32_t *ob, uint32_t *oh, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = __uint_as_float(x[i]);
  bf16x2 v = {};
  v = __builtin_amdgcn_cvt_sr_bf16_f32(v, f, r[i], false);
  ob[i] = __builtin_bit_cast(uint32_t, v);
  half2_ h = {};
  h = __builtin_amdgcn_cvt_sr_f16_f32(h, f, r[i], false);
  oh[i] = __builtin_bit_cast(uint32_t, h);
}

static uint64_t sm(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main(int argc, char **argv) {
  const char *path = argc > 1 ? argv[1] : "sr_hw.bin";
  std::vector<uint32_t> xs = {0u, 0x80000000u, 1u, 0x80000001u, 0x007fffffu, 0x00400000u, 0x00008000u,
                              0x0000ffffu, 0x00010000u, 0x00800000u, 0x80800000u, 0x3f800000u,
                              0xbf800000u, 0x7f7fffffu, 0xff7fffffu, 0x7f7f0000u, 0x7f7f8000u,
                              0x7f800000u, 0xff800000u, 0x7fc00000u, 0x7f800001u, 0x7fffffffu,
                              0xffffffffu, 0x7fbfffffu, 0xff800001u, 0x38800000u, 0x387fffffu,
                              0x33800000u, 0x477fe000u, 0x477fefffu, 0x477ff000u, 0x47800000u,
                              0x38000000u, 0x33000000u, 0x337fffffu, 0x3c00abcdu, 0x3d4ccccdu};
  const uint32_t lows[] = {0x0000u, 0x0001u, 0x1fffu, 0x2000u, 0x7fffu, 0x8000u, 0x8001u, 0xfffeu, 0xffffu};
  for (uint32_t hi : {0x3f80u, 0xbf80u, 0x3c23u, 0x0001u, 0x0080u, 0x7f7fu, 0x4780u})
    for (uint32_t lo : lows) xs.push_back((hi << 16) | lo);
  std::vector<uint32_t> rs = {0u, 1u, 0x1000u, 0x1fffu, 0x2000u, 0x7fffu, 0x8000u, 0xffffu, 0x10000u,
                              0x1ffffu, 0xffff0000u, 0xffffffffu, 0x80000000u, 0x12345678u, 0xdeadbeefu,
                              0x00008000u, 0x80008000u, 0x0000fffeu, 0x00010001u};
  std::vector<uint32_t> X, R;
  for (uint32_t x : xs)
    for (uint32_t r : rs) {
      X.push_back(x);
      R.push_back(r);
    }
  uint64_t s = 12345;
  for (int i = 0; i < (1 << 20); ++i) {
    uint64_t a = sm(s);
    uint32_t x = (uint32_t)a;
    if (i & 1) x = (x & 0x807fffffu) | ((uint32_t)(100 + (a >> 40) % 56) << 23);  // |x| in [2^-27, 2^28]
    X.push_back(x);
    R.push_back((uint32_t)(a >> 32));
  }
  const int n = (int)X.size();
  uint32_t *dx, *dr, *db, *dh;
  if (hipMalloc(&dx, 4 * (size_t)n) || hipMalloc(&dr, 4 * (size_t)n) || hipMalloc(&db, 4 * (size_t)n) ||
      hipMalloc(&dh, 4 * (size_t)n)) {
    fprintf(stderr, "alloc failed\n");
    return 1;
  }
  hipMemcpy(dx, X.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
  hipMemcpy(dr, R.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(probe, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, db, dh, n);
  if (hipDeviceSynchronize() != hipSuccess) {
    fprintf(stderr, "kernel failed\n");
    return 1;
  }
  std::vector<uint32_t> B(n), H(n);
  hipMemcpy(B.data(), db, 4 * (size_t)n, hipMemcpyDeviceToHost);
  hipMemcpy(H.data(), dh, 4 * (size_t)n, hipMemcpyDeviceToHost);
  FILE *fo = fopen(path, "wb");
  fwrite(&n, 4, 1, fo);
  fwrite(X.data(), 4, n, fo);
  fwrite(R.data(), 4, n, fo);
  fwrite(B.data(), 4, n, fo);
  fwrite(H.data(), 4, n, fo);
  fclose(fo);
  printf("probe: %d pairs written to %s\n", n, path);
  return 0;
}
