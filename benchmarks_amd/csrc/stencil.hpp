// 3D 7-point diffusion step kernels for gfx950 (CDNA4, wave64).
//
// One template covers the plain step (T = 1) and temporal blocking (T > 1
// fused steps per launch) for FP32, FP64 and BF16/FP16 storage.
//
// Work decomposition
//   * A workgroup owns an xy tile and a z-chunk.  The tile is TILE_X = 64 lanes
//     x G groups x VEC cells wide (VEC cells = one 16-, 8- or 4-byte vector) and
//     TILE_Y = NW waves x RPW rows tall.  Lane l of group g holds cells
//     x0 + g*64*VEC + l*VEC + [0, VEC) of each of its wave's RPW consecutive
//     rows, so every HBM access is a coalesced 128-bit (64-, 32-bit) access; the
//     4-byte vectors make 64-wide FP32 and 128-wide 16-bit tiles for narrow
//     grids.
//   * The workgroup streams z: iteration i requests plane i + PF (PF = 1..3
//     planes in flight) into registers.  Stage s (1..T) then computes time
//     level s at plane i - s from level s-1 planes (i-s-1, i-s, i-s+1), all
//     held in registers; plane q of a level lives in register slot q mod NS
//     (NS = PF + 3), and the z-loop is unrolled by NS so slots are
//     compile-time and never moved.
//   * x neighbours: inside a lane from registers, across lanes by DPP
//     wave_shr:1 / wave_shl:1 (group seams via readlane).  y neighbours: from
//     the wave's own rows in registers; a wave's first/last rows are exchanged
//     with the adjacent waves through LDS (double-buffered with one barrier
//     per plane when that fits in 64 KB, else two barriers).  Tile edges use
//     the cell itself, which is exactly the clamped boundary at a domain
//     edge; at interior tile edges the resulting invalid cells are never
//     stored (tiles overlap by the halo, see Geometry), and the last stage
//     skips the arithmetic of the halo rows (their stores are dropped).
//   * Clamp selects (x = nx-1, y = ny-1, z = 0 / nz-1) are only evaluated in
//     the EDGE variant of a stage, used for tiles that contain the grid's last
//     column or row beyond their own last cell, and for the first/last plane.
//   * Arithmetic: EXACT evaluates cc*c + cw*w + ce*e + cs*s + cn*n + cb*b +
//     ct*t left to right with separately rounded operations (the reference's
//     order, diffusion3d/baseline.cc:39-40; the file is built with
//     -ffp-contract=off), so results are bitwise equal to the reference CPU
//     path.  FMA uses fma(cc,c,cw*w) followed by an fma chain.
//   * Reduced-precision storage rounds every time level (also the fused
//     intermediate ones) to storage precision, so T fused steps are bitwise
//     equal to T single steps for every dtype.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace d3h {

enum StoreKind : int { kF32 = 0, kF64 = 1, kBF16 = 2, kF16 = 3 };

template <int S> struct Traits;
template <> struct Traits<kF32> { using T = float; using C = float; };
template <> struct Traits<kF64> { using T = double; using C = double; };
template <> struct Traits<kBF16> { using T = uint16_t; using C = float; };
template <> struct Traits<kF16> { using T = uint16_t; using C = float; };

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two FP32 values to packed BF16, round to nearest even: gfx950's
// v_cvt_pk_bf16_f32.  It equals to_bf16 below on every one of the 2^32 input
// patterns (NaNs quieted by setting bit 6, subnormals rounded, not flushed),
// checked exhaustively on an MI355X.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float lo, float hi) {
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Buffer descriptor word 3 for untyped (raw) buffer access on the gfx9
// family: DATA_FORMAT = 32 bit, so range checking is per dword.
constexpr int kRsrcWord3 = 0x00020000;
// Cache policy bits of buffer instructions on gfx94x/gfx950: bit 1 = nt.
#ifndef D3H_LOAD_AUX
#define D3H_LOAD_AUX 0
#endif
constexpr int kLoadAux = D3H_LOAD_AUX;
constexpr int kStoreAux = 2;
// Offset that lies beyond every descriptor (planes are < 2 GiB).
constexpr int kOutOfRange = 0x7ffffff0;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, bytes, kRsrcWord3);
}

// Launch geometry, computed on the host (make_geometry in runtime.hip).
//   Tiles overlap: tile (tx,ty,tz) stores x in [tx*ox, tx*ox+ox), rows
//   [ty*oy, ty*oy+oy), planes [tz*oz, tz*oz+oz) (clipped to the grid) and loads
//   a window that extends hx cells / T rows / T planes beyond that on every
//   side that is not a domain boundary (the last tile's window is moved back
//   to end at the grid's edge where it fits, see TileRunner).  pitch is the padded row length
//   (multiple of 16 bytes); plane = pitch * ny.
struct Geometry {
  int nx, ny, nz, pitch;
  long long plane;
  int ntx, nty, nzc;
  int ox, oy, oz, hx;
  // Stochastic rounding of BF16/FP16 storage (sr = 1): every stored time level is
  // rounded with a hash of (seed, cell, step); step0 is the number of steps
  // the field had taken before this launch (a batched launch adds launch * T).
  int sr;
  unsigned seed;
  int step0;
  // 2D fields streamed RPW rows per "plane" (stencil2d_kernel): the field's
  // row count; plane q holds rows [q*RPW, q*RPW + RPW) (the last one clipped).
  // 0 for every other kernel.
  int rows2d;
  // Plane schedule of depth-1 tiles (0: none): a workgroup starts its z
  // iteration j (counted from its first) no earlier than its start time plus
  // j * pace / 256 ticks of the device's real-time counter.
  int pace;
};

template <typename C> struct Coef { C ce, cw, cn, cs, ct, cb, cc; };

// ---- storage conversion (round-to-nearest-even) ----------------------------
__device__ __forceinline__ uint32_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float f16_bits_to_f32(uint32_t h) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(h & 0xffffu));
}
__device__ __forceinline__ uint32_t to_f16(float f) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
}

// VEC storage elements <-> NWORDS 32-bit words.
template <int S, int VEC>
struct Pack {
  using C = typename Traits<S>::C;
  static constexpr int kBytes = VEC * (int)sizeof(typename Traits<S>::T);
  static constexpr int kWords = kBytes / 4;
  static_assert(kBytes == 4 || kBytes == 8 || kBytes == 16, "lane vectors are 4, 8 or 16 bytes");

  __device__ static __forceinline__ void unpack(const uint32_t *w, C *v) {
    if constexpr (S == kF32) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) v[i] = __uint_as_float(w[i]);
    } else if constexpr (S == kF64) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) v[i] = __hiloint2double((int)w[2 * i + 1], (int)w[2 * i]);
    } else if constexpr (S == kBF16) {
#pragma unroll
      for (int i = 0; i < kWords; ++i) {
        v[2 * i] = __uint_as_float(w[i] << 16);
        v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
      }
    } else {
#pragma unroll
      for (int i = 0; i < kWords; ++i) {
        v[2 * i] = f16_bits_to_f32(w[i]);
        v[2 * i + 1] = f16_bits_to_f32(w[i] >> 16);
      }
    }
  }
  __device__ static __forceinline__ void pack(const C *v, uint32_t *w) {
    if constexpr (S == kF32) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) w[i] = __float_as_uint(v[i]);
    } else if constexpr (S == kF64) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) {
        w[2 * i] = (uint32_t)__double2loint(v[i]);
        w[2 * i + 1] = (uint32_t)__double2hiint(v[i]);
      }
    } else if constexpr (S == kBF16) {
#pragma unroll
      for (int i = 0; i < kWords; ++i) w[i] = cvt_pk_bf16(v[2 * i], v[2 * i + 1]);
    } else {
#pragma unroll
      for (int i = 0; i < kWords; ++i) w[i] = to_f16(v[2 * i]) | (to_f16(v[2 * i + 1]) << 16);
    }
  }
  // Unconditional buffer accesses: an offset at or beyond the descriptor's
  // size reads zeros / drops the store, so no lane or plane needs a branch
  // (branch-free memory keeps the compiler's wait counts exact, which lets
  // the next plane's loads stay in flight during the current plane's math).
  __device__ static __forceinline__ void load(__amdgpu_buffer_rsrc_t rs, int off, C *v) {
    uint32_t w[kWords];
    if constexpr (kBytes == 16) {
      const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, kLoadAux);
      w[0] = u.x; w[1] = u.y; w[2] = u.z; w[3] = u.w;
    } else if constexpr (kBytes == 8) {
      const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, kLoadAux);
      w[0] = u.x; w[1] = u.y;
    } else {
      w[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, kLoadAux);
    }
    unpack(w, v);
  }
  __device__ static __forceinline__ void store(__amdgpu_buffer_rsrc_t rs, int off, const C *v) {
    uint32_t w[kWords];
    pack(v, w);
    store_words(rs, off, w);
  }
  // Store already packed words (kWords of them).
  __device__ static __forceinline__ void store_words(__amdgpu_buffer_rsrc_t rs, int off,
                                                     const uint32_t *w) {
    if constexpr (kBytes == 16) {
      u32x4 u = {w[0], w[1], w[2], w[3]};
      __builtin_amdgcn_raw_buffer_store_b128(u, rs, off, 0, kStoreAux);
    } else if constexpr (kBytes == 8) {
      u32x2 u = {w[0], w[1]};
      __builtin_amdgcn_raw_buffer_store_b64(u, rs, off, 0, kStoreAux);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(w[0], rs, off, 0, kStoreAux);
    }
  }
};

// Deterministic stochastic rounding of 16-bit storage.  Cell (x, y, z) at
// step t takes 16 random bits r from
//   base = lowbias32(row * 0x9E3779B1 + t * 0x85EBCA77 + seed),  row = z*ny + y
//   h    = base + (x / 2) * 0x9E3779B9
// (lowbias32: C. Wellons' hash-prospector integer mixer): the high half of h
// for even x, the low half for odd x.  base is uniform per wave (rows belong
// to waves), so it is computed on the scalar unit; a cell pair costs one add.
// gfx950 rounds with v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32, which take the
// high 16 bits of their random operand: h for the even cell, h << 16 for the
// odd one.  BF16: r is added below the mantissa and the sum truncated (inf and
// nan round to nearest).  FP16: |x| + (r >> 3) FP16 ulps / 2^13 truncated to
// the FP16 grid, subnormals included.  Unbiased, and reproducible bitwise on
// the host (the oracle restates these rules; a probe over special and random
// operands pinned them on an MI355X).
constexpr uint32_t kSrRowMul = 0x9E3779B1u;
constexpr uint32_t kSrStepMul = 0x85EBCA77u;
constexpr uint32_t kSrPairWeyl = 0x9E3779B9u;

__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// A value the program never reads, for the results of rows whose stores are
// dropped: an empty asm statement defines it, so no instruction sets the
// register.  The operand u (a value live at that point) keeps each statement
// distinct, so they are not merged into one value copied around.
template <class V, class U>
__device__ __forceinline__ V unset_like(U u) {
  V v;
  asm("; unset %0 (%1)" : "=v"(v) : "v"(u));
  return v;
}

// Two FP32 values (even and odd cell of one pair) stochastically rounded to
// packed 16-bit storage with the pair's random word h.  Each conversion writes
// one half of the word and keeps the other, so the word starts unset (no
// initialising move; the register can be the store's own).
template <int S>
__device__ __forceinline__ uint32_t sr_pack_pair(float lo, float hi, uint32_t h) {
  if constexpr (S == kBF16) {
    bf16x2 v = unset_like<bf16x2>(lo);
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, lo, h, false);
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, hi, h << 16, true);
    return __builtin_bit_cast(uint32_t, v);
  } else {
    f16x2 v = unset_like<f16x2>(lo);
    v = __builtin_amdgcn_cvt_sr_f16_f32(v, lo, h, false);
    v = __builtin_amdgcn_cvt_sr_f16_f32(v, hi, h << 16, true);
    return __builtin_bit_cast(uint32_t, v);
  }
}

// ---- cross-lane moves -------------------------------------------------------
// DPP wave_shr:1 (lane l <- lane l-1) and wave_shl:1 (lane l <- lane l+1);
// the lane without a source keeps `old`.
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL>
__device__ __forceinline__ float dpp(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src),
                                                    CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double old, double src) {
  int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(src), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(src), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// coef * (the neighbour lane's src) in one instruction: v_mul_f32 with a DPP
// wave_shr:1 / wave_shl:1 source.  The lane without a source keeps `old` (the
// instruction's destination is tied to it and no bound_ctrl is set), so the
// caller passes old = coef * (its own clamp value) and gets the product the
// separate move, DPP move and multiply gave, bitwise (one rounding of the
// same product).  coef_v is a VGPR (VOP2 DPP takes no scalar second operand).
// The s_nop gives the 2 wait states a DPP read of a VGPR needs after a VALU
// write of it, since the compiler does not look inside the statement.
template <int CTRL>
__device__ __forceinline__ float mul_dpp(float old, float src, float coef_v) {
  float d = old;
  if constexpr (CTRL == kWaveShr1)
    asm("s_nop 1\n\tv_mul_f32_dpp %0, %1, %2 wave_shr:1 row_mask:0xf bank_mask:0xf"
        : "+v"(d) : "v"(src), "v"(coef_v));
  else
    asm("s_nop 1\n\tv_mul_f32_dpp %0, %1, %2 wave_shl:1 row_mask:0xf bank_mask:0xf"
        : "+v"(d) : "v"(src), "v"(coef_v));
  return d;
}
// A copy of a uniform value that stays in a VGPR (the empty statement hides
// that it is uniform, so it is not rematerialised from the scalar per use).
__device__ __forceinline__ float in_vgpr(float v) {
  float r;
  asm("; vgpr copy %0" : "=v"(r) : "0"(v));
  return r;
}
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The 7-point update; argument order follows the reference expression
// (s = y+1, n = y-1, b = z-1, t = z+1).
template <typename C, bool FMA>
__device__ __forceinline__ C update(const Coef<C> &k, C c, C w, C e, C s, C n, C b, C t) {
  if constexpr (FMA) {
    C r = fma_(k.cc, c, k.cw * w);
    r = fma_(k.ce, e, r);
    r = fma_(k.cs, s, r);
    r = fma_(k.cn, n, r);
    r = fma_(k.cb, b, r);
    return fma_(k.ct, t, r);
  } else {
    return k.cc * c + k.cw * w + k.ce * e + k.cs * s + k.cn * n + k.cb * b + k.ct * t;
  }
}

// The 2D 5-point update, the diffusion/ family's term order
// (diffusion/diffusion_openmp.cc:59-82 on a row-reversed field: s = device row
// y + 1 = host row y - 1 first).
template <typename C, bool FMA>
__device__ __forceinline__ C update5(const Coef<C> &k, C c, C w, C e, C s, C n) {
  if constexpr (FMA) {
    C r = fma_(k.cc, c, k.cw * w);
    r = fma_(k.ce, e, r);
    r = fma_(k.cs, s, r);
    return fma_(k.cn, n, r);
  } else {
    return k.cc * c + k.cw * w + k.ce * e + k.cs * s + k.cn * n;
  }
}

template <int N>
constexpr int modn(int v) { return ((v % N) + N) % N; }

// Exchange buffer between the waves of a workgroup: [buf][level][wave][side][v][lane].
// Double-buffered (one barrier per plane) when both buffers fit in 64 KB,
// otherwise single-buffered with a second barrier before each publish.
template <typename C, int T, int NV, int NW>
struct XchLayout {
  static constexpr int kPerBuf = T * NW * 2 * NV * 64;
  static constexpr int kBufs = (2 * kPerBuf * (int)sizeof(C) <= 65536) ? 2 : 1;
  static constexpr int kElems = (NW > 1) ? kBufs * kPerBuf : 1;
  __device__ static constexpr int idx(int buf, int lvl, int wave, int side, int v, int lane) {
    return ((((buf * T + lvl) * NW + wave) * 2 + side) * NV + v) * 64 + lane;
  }
};

// PF = prefetch distance in planes: plane i + PF is requested while plane i
// is computed.  Each level keeps NS = PF + 3 register slots (plane q in slot
// q mod NS) and the z loop is unrolled by NS.
//
// PF = 0 (fused depths T >= 2 only) is the three-slot ring: every level,
// level 0 included, keeps just the 3 planes its next stage reads, and plane
// i + 1 is requested right after stage 1 of iteration i, into the slot of
// plane i - 2 that stage 1 has just read for the last time; stages 2..T then
// hide its latency.  That frees T * RPW * NV registers per lane against
// PF = 1, enough for taller tiles at the same occupancy.
//
// W2D (one wave, NW = 1): the 2D 5-point operator on an nx x ny field whose
// "planes" are blocks of RPW consecutive rows.  Row r's y neighbours are rows
// r +- 1 of the same plane, or, across a plane boundary, the last row of the
// previous plane / the first row of the next one (register slots of the
// adjacent planes); there are no z terms.
//
// REV (2D mode only): the chunk streams its row blocks from the last to the
// first.  Stream index q is then row block nz - 1 - q (the chunk's ranges are
// mirrored), and a block's neighbour blocks swap register slots.  Odd chunks
// stream backwards, so the boundary blocks two adjacent chunks both read are
// read by both at about the same time (at the end of both, or at the start of
// both) and the second read is served from L2.
template <int S, int VEC, int T, int G, int RPW, int NW, bool FMA, int PF, bool W2D = false,
          bool REV = false>
struct TileRunner {
  static_assert(!W2D || NW == 1, "2D row blocks are streamed by one wave");
  static_assert(!REV || W2D, "only 2D row blocks stream backwards");
  using ST = typename Traits<S>::T;
  using C = typename Traits<S>::C;
  static constexpr int NV = G * VEC;
  using X = XchLayout<C, T, NV, NW>;
  using P = Pack<S, VEC>;
  static constexpr int GSTRIDE = 64 * VEC;  // cells between a lane's groups
  static constexpr int TILE_X = G * GSTRIDE;
  static constexpr int TILE_Y = NW * RPW;
  static constexpr int NS = PF + 3;
  // Halo planes of a z-chunk: T (one plane per fused step), or in the 2D mode,
  // where a step reaches one row and a "plane" holds RPW rows, ceil(T / RPW).
  static constexpr int kZHalo = W2D ? (T + RPW - 1) / RPW : T;
  static_assert(PF > 0 || (T >= 2 && !W2D), "the three-slot ring needs fused depths");
  // A tile taller than its 2T halo rows has halo rows whose last-stage values
  // are never stored; the last stage skips their arithmetic.  (FP64 tiles,
  // whose 16 bytes per cell keep them far from the VALU limit, measured 1.4 %
  // slower with the skip, so they compute every row.)
  static constexpr bool kSkipHaloRows = TILE_Y > 2 * T && S != kF64;
  // Depths 2 and 4 also skip, at stages 1..T-1, the rows whose level no
  // stored row reads: row y at stage s feeds the stored rows only within
  // T - s rows of them (the tile's first and last waves hold such rows).
  // Same-box A/B: 768^3 depth 4 -2 %, depth 2 -0.5..-1.7 %, the batched
  // 512-grid FP32 sweep (depth 2) +5 % solves/s; depth 8 measured 16 % slower
  // with it and computes every row.
  static constexpr bool kSkipEarlyRows = kSkipHaloRows && T >= 2 && T <= 4 && !W2D;
  // The first / last wave of a workgroup takes the cell itself as its outer
  // y neighbour (the clamp at the tile's edge).  At T <= 2 it reads that row
  // back from the exchange buffer (where it published it), so the choice is an
  // address and each stage one basic block: batched 512-grid sweeps measured
  // 1 % (FP32, depth 2) and 5 % (BF16) faster, 768^3 depth 2 2 % faster.  The
  // deep tiles keep the branch: depth 4 measured 5 % slower with the reads.
  static constexpr bool kEdgeRowsFromLds = T <= 2;
  // EXACT FP32 arithmetic (also for 16-bit storage) multiplies the x terms of
  // a lane's edge cells in the DPP instruction that fetches the neighbour
  // (mul_dpp): 2 VALU instructions per row and stage fewer than a move of the
  // clamp value, a DPP move and a multiply for each side.  The EDGE variant
  // keeps the moves (its clamp selects need the neighbour's value).
  static constexpr bool kXMul = !FMA && S != kF64;

  // lv[level][slot][row][cell]: level 0 is the loaded input.
  C lv[T][NS][RPW][NV];

  const ST *__restrict__ src;
  ST *__restrict__ dst;
  const Geometry &geo;
  const Coef<C> &k;
  C *xch;
  int lane, wave;
  int xlane;            // x of this lane's cell 0 in group 0
  int yrow0;            // y of this wave's row 0
  int zl0, zl1, zo0, zo1;
  bool edge_xy;         // tile holds the last column/row strictly inside it
  int plane_bytes;
  int row_bytes;         // W2D: bytes of one row (pitch)
  int lastr;             // W2D: row of the last plane holding the field's last row
  int ldoff[RPW][G];   // byte offset in a plane, or kOutOfRange (loads zeros)
  int stoff[RPW][G];   // byte offset in a plane, or kOutOfRange (store dropped)
  bool rowlast[RPW];
  int rowhalo[RPW];     // row is not stored: 0 / 1, uniform per wave (a scalar register)
  int rowskip;          // bit (s - 1) * RPW + r: stage s need not compute row r (uniform)
  uint32_t sr_lane[G];  // 16-bit storage: Weyl offset of the lane's first x pair per group
  float cwv, cev;       // cw and ce in VGPRs (mul_dpp's second operand)
  uint64_t t_start;     // depth 1: real-time counter at the workgroup's start
  int i_first;          // depth 1: first z iteration (the plane schedule's origin)

  __device__ __forceinline__ TileRunner(const ST *s_, ST *d_, const Geometry &g, const Coef<C> &k_,
                                        C *x_, int tx, int ty, int tz)
      : src(s_), dst(d_), geo(g), k(k_), xch(x_) {
    if constexpr (T == 1) t_start = __builtin_amdgcn_s_memrealtime();
    lane = __lane_id();
    wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int xo0 = tx * geo.ox;
    const int xo1 = min(xo0 + geo.ox, geo.nx);
    const int xo1s = (xo1 == geo.nx) ? geo.pitch : xo1;  // row padding may be overwritten
    int xl0 = (geo.ntx == 1) ? 0 : max(xo0 - geo.hx, 0);
    const int yo0 = ty * geo.oy;
    const int yo1 = min(yo0 + geo.oy, geo.ny);
    int yl0 = (geo.nty == 1) ? 0 : max(yo0 - T, 0);
    // A window that would reach past the grid's last column (row) is moved
    // back to end exactly there, when the grid is at least a tile wide (tall)
    // and the shifted start stays on a lane-vector boundary: its stored cells
    // still lie T or more cells inside the window, and the last lane's (last
    // wave's) own cell is then the clamped neighbour, so the tile needs no
    // clamp selects.
    if (xl0 + TILE_X > geo.nx && geo.nx >= TILE_X && (geo.nx - TILE_X) % VEC == 0)
      xl0 = geo.nx - TILE_X;
    if (yl0 + TILE_Y > geo.ny && geo.ny >= TILE_Y) yl0 = geo.ny - TILE_Y;
    zo0 = tz * geo.oz;
    zo1 = min(zo0 + geo.oz, geo.nz);
    if constexpr (REV) {
      const int r0 = zo0;
      zo0 = geo.nz - zo1;
      zo1 = geo.nz - r0;
    }
    zl0 = max(zo0 - kZHalo, 0);
    zl1 = min(zo1 + kZHalo, geo.nz);
    edge_xy = (xl0 + TILE_X > geo.nx) || (yl0 + TILE_Y > geo.ny);
    xlane = xl0 + lane * VEC;
    yrow0 = yl0 + wave * RPW;
    plane_bytes = (int)(geo.plane * (long long)sizeof(ST));
    lastr = W2D ? geo.rows2d - 1 - (geo.nz - 1) * RPW : RPW - 1;
    row_bytes = geo.pitch * (int)sizeof(ST);
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int y = yl0 + wave * RPW + r;
      const bool rowin = y < geo.ny;
      const bool rowst = y >= yo0 && y < yo1;
      rowlast[r] = y == geo.ny - 1;
      rowhalo[r] = __builtin_amdgcn_readfirstlane(rowst ? 0 : 1);

#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int xv = xlane + g * GSTRIDE;
        const int off = (y * geo.pitch + xv) * (int)sizeof(ST);
        ldoff[r][g] = (rowin && xv < geo.pitch) ? off : kOutOfRange;
        stoff[r][g] = (rowst && xv >= xo0 && xv < xo1s) ? off : kOutOfRange;
      }
    }
    if constexpr (kXMul) {
      cwv = in_vgpr((float)k.cw);
      cev = in_vgpr((float)k.ce);
    }
    if constexpr (kSkipEarlyRows) {
      int m = 0;
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        const int y = yl0 + wave * RPW + r;
#pragma unroll
        for (int st = 1; st < T; ++st) {
          const bool need = y >= yo0 - (T - st) && y < yo1 + (T - st) && y < geo.ny;
          if (!need) m |= 1 << ((st - 1) * RPW + r);
        }
      }
      rowskip = __builtin_amdgcn_readfirstlane(m);
    }
    if constexpr (S == kBF16 || S == kF16) {
      // xlane is a multiple of VEC (even), so a lane's x pairs are its own
#pragma unroll
      for (int g = 0; g < G; ++g) sr_lane[g] = ((uint32_t)(xlane + g * GSTRIDE) >> 1) * kSrPairWeyl;
    }
#pragma unroll
    for (int l = 0; l < T; ++l)
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
          for (int v = 0; v < NV; ++v) lv[l][s][r][v] = C(0);
  }

  // The field's plane (2D: row block) at stream index q.
  __device__ __forceinline__ int real(int q) const { return REV ? geo.nz - 1 - q : q; }

  // Bytes of plane q that belong to the field: the whole plane, except the
  // clipped last row block of a 2D field (rows beyond it read zeros and their
  // stores are dropped).
  __device__ __forceinline__ int valid_bytes(int q) const {
    if constexpr (W2D) return min(RPW, geo.rows2d - q * RPW) * row_bytes;
    else return plane_bytes;
  }

  // Load plane q into level-0 slot SL; planes outside [zl0, zl1) read zeros.
  template <int SL>
  __device__ __forceinline__ void load_plane(int q) {
    const bool in = q >= zl0 && q < zl1;
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(src + (long long)real(q) * geo.plane, in ? valid_bytes(real(q)) : 0);
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int g = 0; g < G; ++g) P::load(rs, ldoff[r][g], &lv[0][SL][r][g * VEC]);
  }

  // Store plane q; planes outside [zo0, zo1) are dropped.
  __device__ __forceinline__ void store_plane(int q, C (&out)[RPW][NV]) {
    const bool in = q >= zo0 && q < zo1;
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(dst + (long long)real(q) * geo.plane, in ? valid_bytes(real(q)) : 0);
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int g = 0; g < G; ++g) P::store(rs, stoff[r][g], &out[r][g * VEC]);
  }

  // Store plane q from packed 16-bit words (NV / 2 per row); planes outside
  // [zo0, zo1) are dropped.
  __device__ __forceinline__ void store_plane_words(int q, const uint32_t (&w)[RPW][NV / 2]) {
    const bool in = q >= zo0 && q < zo1;
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(dst + (long long)real(q) * geo.plane, in ? valid_bytes(real(q)) : 0);
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int g = 0; g < G; ++g) P::store_words(rs, stoff[r][g], &w[r][g * VEC / 2]);
  }

  // Publish the first/last row of every level's centre plane for iteration
  // phase PH (plane i - s of level s - 1 for stage s).
  template <int PH>
  __device__ __forceinline__ void publish(int buf) {
    if constexpr (NW > 1) {
#pragma unroll
      for (int s = 1; s <= T; ++s) {
        const int sc = modn<NS>(PH - s);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          xch[X::idx(buf, s - 1, wave, 0, v, lane)] = lv[s - 1][sc][0][v];
          xch[X::idx(buf, s - 1, wave, 1, v, lane)] = lv[s - 1][sc][RPW - 1][v];
        }
      }
    }
  }

  template <int PH, int s, bool EDGE>
  __device__ __forceinline__ void stage_impl(int p, int buf) {
    constexpr int sc = modn<NS>(PH - s), sb = modn<NS>(PH - s - 1), st = modn<NS>(PH - s + 1);
    // the field's first / last plane (row block)
    const bool zb = (real(p) == 0), zt = (real(p) == geo.nz - 1);
    const int wn = max(wave - 1, 0), ws = min(wave + 1, NW - 1);
    constexpr bool k16 = (S == kBF16 || S == kF16);
    C out[RPW][NV];
    // 16-bit storage: every row is rounded to packed storage words (nearest
    // even or stochastic; the branch on geo.sr only touches registers), which
    // are then either unpacked as the next level or stored.  One store site,
    // so the memory instructions of an iteration do not depend on the mode.
    uint32_t wv[RPW][k16 ? NV / 2 : 1];
    const uint32_t sbase = (uint32_t)(geo.step0 + s - 1) * kSrStepMul + geo.seed;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const C *cv = lv[s - 1][sc][r];
      const bool skip = (s == T) ? (kSkipHaloRows && rowhalo[r] != 0)
                                 : (kSkipEarlyRows && ((rowskip >> ((s - 1) * RPW + r)) & 1) != 0);
      if (skip) {
        // uniform branch: the store of this row is dropped anyway (last
        // stage), or no row that will be stored reads this row's level
        if constexpr (k16) {
#pragma unroll
          for (int w = 0; w < NV / 2; ++w) wv[r][w] = unset_like<uint32_t>(cv[2 * w]);
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) out[r][v] = unset_like<C>(cv[v]);
        }
        continue;
      }
      // rows above / below this wave's first / last row
      C nrow[NV], srow[NV];
      if (W2D && r == 0) {
        // the previous block's last row (its value at a field edge is
        // replaced by the cell itself below); streaming backwards, the block
        // above is the next one in the stream
        constexpr int sn = REV ? st : sb;
#pragma unroll
        for (int v = 0; v < NV; ++v) nrow[v] = lv[s - 1][sn][RPW - 1][v];
      } else if (r == 0) {
        if constexpr (NW > 1 && kEdgeRowsFromLds) {
          // the previous wave's last row; wave 0 reads back its own first row
          // (the cell itself: the clamp at the tile's top edge)
          const int nside = wave > 0 ? 1 : 0;
#pragma unroll
          for (int v = 0; v < NV; ++v) nrow[v] = xch[X::idx(buf, s - 1, wn, nside, v, lane)];
        } else if constexpr (NW > 1) {
          if (wave > 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) nrow[v] = xch[X::idx(buf, s - 1, wn, 1, v, lane)];
          } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) nrow[v] = cv[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) nrow[v] = cv[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) nrow[v] = lv[s - 1][sc][r - 1][v];
      }
      if (W2D && r == RPW - 1) {
        // the next block's first row (streaming backwards: the previous one
        // in the stream)
        constexpr int ss = REV ? sb : st;
#pragma unroll
        for (int v = 0; v < NV; ++v) srow[v] = lv[s - 1][ss][0][v];
      } else if (r == RPW - 1) {
        if constexpr (NW > 1 && kEdgeRowsFromLds) {
          // the next wave's first row; the last wave reads back its own last row
          const int sside = wave < NW - 1 ? 0 : 1;
#pragma unroll
          for (int v = 0; v < NV; ++v) srow[v] = xch[X::idx(buf, s - 1, ws, sside, v, lane)];
        } else if constexpr (NW > 1) {
          if (wave < NW - 1) {
#pragma unroll
            for (int v = 0; v < NV; ++v) srow[v] = xch[X::idx(buf, s - 1, ws, 0, v, lane)];
          } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) srow[v] = cv[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) srow[v] = cv[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) srow[v] = lv[s - 1][sc][r + 1][v];
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        // x neighbours across lanes (group seams through readlane)
        const C wold = (g > 0) ? readlane(cv[(g - 1) * VEC + VEC - 1], 63) : cv[0];
        const C eold = (g < G - 1) ? readlane(cv[(g + 1) * VEC], 0) : cv[g * VEC + VEC - 1];
        if constexpr (kXMul && !EDGE) {
          // The reference's sum, cc c + cw w + ce e + ..., accumulated term by
          // term; the x terms of the lane's first and last cells are multiplied
          // as they cross lanes (the lane without a neighbour keeps coef * its
          // old value).  The in-lane terms come first: for group 0 the DPP's
          // old values cw * c0 and ce * c_last are also cell 1's w term and
          // cell VEC-2's e term, and are then dead when the DPP overwrites
          // them (no copies).
          const int g0 = g * VEC;
          C acc[VEC];
#pragma unroll
          for (int v = 1; v < VEC; ++v) acc[v] = k.cc * cv[g0 + v] + k.cw * cv[g0 + v - 1];
          const C pw = mul_dpp<kWaveShr1>(k.cw * wold, cv[g0 + VEC - 1], cwv);
          acc[0] = k.cc * cv[g0] + pw;
#pragma unroll
          for (int v = 0; v < VEC - 1; ++v) acc[v] = acc[v] + k.ce * cv[g0 + v + 1];
          const C pe = mul_dpp<kWaveShl1>(k.ce * eold, cv[g0], cev);
          acc[VEC - 1] = acc[VEC - 1] + pe;
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            const int gv = g0 + v;
            if constexpr (W2D)
              out[r][gv] = acc[v] + k.cs * srow[gv] + k.cn * nrow[gv];
            else
              out[r][gv] = acc[v] + k.cs * srow[gv] + k.cn * nrow[gv] + k.cb * lv[s - 1][sb][r][gv] +
                           k.ct * lv[s - 1][st][r][gv];
          }
          continue;
        }
        const C wl = dpp<kWaveShr1>(wold, cv[g * VEC + VEC - 1]);
        const C er = dpp<kWaveShl1>(eold, cv[g * VEC]);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int gv = g * VEC + v;
          const C c = cv[gv];
          const C w = (v > 0) ? cv[gv - 1] : wl;
          C e = (v < VEC - 1) ? cv[gv + 1] : er;
          C so = srow[gv];
          if constexpr (W2D) {
            C no = nrow[gv];
            if constexpr (EDGE) {
              if (xlane + g * GSTRIDE + v == geo.nx - 1) e = c;
              if (zb && r == 0) no = c;         // field row 0
              if (zt && r == lastr) so = c;     // the field's last row
            }
            out[r][gv] = update5<C, FMA>(k, c, w, e, so, no);
          } else {
            C b = lv[s - 1][sb][r][gv];
            C t = lv[s - 1][st][r][gv];
            if constexpr (EDGE) {
              if (xlane + g * GSTRIDE + v == geo.nx - 1) e = c;
              if (rowlast[r]) so = c;
              if (zb) b = c;
              if (zt) t = c;
            }
            out[r][gv] = update<C, FMA>(k, c, w, e, so, nrow[gv], b, t);
          }
        }
      }
      if constexpr (k16) {
        if (geo.sr) {
          // uniform per wave: the scalar unit computes it
          const uint32_t row = (uint32_t)real(p) * (uint32_t)geo.ny + (uint32_t)(yrow0 + r);
          const uint32_t base = lowbias32(row * kSrRowMul + sbase);
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const uint32_t hb = base + sr_lane[g];
#pragma unroll
            for (int v = 0; v < VEC; v += 2)
              wv[r][(g * VEC + v) / 2] = sr_pack_pair<S>(out[r][g * VEC + v], out[r][g * VEC + v + 1],
                                                          hb + (uint32_t)(v / 2) * kSrPairWeyl);
          }
        } else {
#pragma unroll
          for (int g = 0; g < G; ++g) P::pack(&out[r][g * VEC], &wv[r][g * VEC / 2]);
        }
      }
    }
    if constexpr (k16) {
      if constexpr (s < T) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
          for (int g = 0; g < G; ++g) P::unpack(&wv[r][g * VEC / 2], &lv[s][sc][r][g * VEC]);
      } else {
        store_plane_words(p, wv);
      }
    } else {
      if constexpr (s < T) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
          for (int v = 0; v < NV; ++v) lv[s][sc][r][v] = out[r][v];
      } else {
        store_plane(p, out);
      }
    }
  }

  // Every stage runs every iteration (planes outside the valid range compute
  // values that are never stored or read by a valid plane), so each path
  // through an iteration issues the same memory instructions.  (Skipping the
  // arithmetic of the planes no stored plane needs, behind a uniform branch
  // that still issued the dropped stores, measured 7-10 % slower on the 512^3
  // BF16 and 4-wave FP32 tiles and changed depths 2 and 4 by under 1 %.)
  template <int PH, int s>
  __device__ __forceinline__ void stage(int i, int buf) {
    const int p = i - s;
    if (edge_xy || p <= 0 || p >= geo.nz - 1) stage_impl<PH, s, true>(p, buf);
    else stage_impl<PH, s, false>(p, buf);
  }

  template <int PH, int s>
  __device__ __forceinline__ void stages(int i, int buf) {
    stage<PH, s>(i, buf);
    if constexpr (s < T) stages<PH, s + 1>(i, buf);
  }

  // Depth 1: hold the workgroup to its plane schedule (geo.pace) before it
  // requests the next plane.  Every workgroup streams its own column of
  // planes; held to one schedule, the streams of all workgroups stay on
  // nearby planes instead of drifting apart over the launch, and HBM serves
  // such compact fronts faster (the autotuner measures whether a schedule
  // helps and how fast it may run).  A workgroup that is behind never waits,
  // and the wait is bounded (4096 sleeps of 64 cycles), so a schedule can
  // only slow a launch down, never stall it.
  __device__ __forceinline__ void pace(int i) {
    if (geo.pace > 0) {
      const uint64_t due = t_start + (((uint64_t)(uint32_t)(i - i_first) * (uint32_t)geo.pace) >> 8);
      for (int n = 0; n < 4096 && __builtin_amdgcn_s_memrealtime() < due; ++n) __builtin_amdgcn_s_sleep(1);
    }
  }

  // Iteration i (i == PH mod NS): request plane i + PF, exchange the wave
  // edge rows, run the stages (PF = 0: request plane i + 1 after stage 1).
  // The exchange buffer alternates with i.
  template <int PH>
  __device__ __forceinline__ void iter(int i) {
    if constexpr (T == 1) pace(i);
    if constexpr (PF > 0) load_plane<modn<NS>(PH + PF)>(i + PF);
    const int buf = (X::kBufs == 2) ? (i & 1) : 0;
    if constexpr (NW > 1) {
      if constexpr (X::kBufs == 1) __syncthreads();  // previous plane's reads are done
      publish<PH>(buf);
      __syncthreads();
    }
    if constexpr (PF > 0) {
      stages<PH, 1>(i, buf);
    } else {
      stage<PH, 1>(i, buf);
      load_plane<modn<NS>(PH + 1)>(i + 1);  // the slot of plane i - 2, read last by stage 1
      stages<PH, 2>(i, buf);
    }
  }

  template <int PH>
  __device__ __forceinline__ void iters(int i) {
    iter<PH>(i + PH);
    if constexpr (PH + 1 < NS) iters<PH + 1>(i);
  }

  __device__ __forceinline__ void run() {
    // the first plane (zl0) is requested at iteration zl0 - PF (PF = 0: at
    // iteration zl0 - 1, after its first stage).  Phases (register slots)
    // count from that iteration; the loop runs whole groups of NS iterations,
    // so a chunk whose iteration count is not a multiple of NS computes up to
    // NS - 1 planes past its end (their stores are dropped).  The host sizes
    // interior chunks to whole groups.
    const int i0 = zl0 - (PF > 0 ? PF : 1);
    i_first = i0;
    const int ilast = zo1 - 1 + T;
    for (int i = i0; i <= ilast; i += NS) iters<0>(i);
  }
};

// Map the hardware workgroup id to a tile so that the workgroups one XCD runs
// (ids xcd, xcd+8, xcd+16, ...) get consecutive tiles, i.e. spatial neighbours
// whose shared halo rows stay in that XCD's L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, xcd = b & 7, slot = b >> 3;
  return xcd * q + min(xcd, r) + slot;
}

template <int S, int VEC, int T, int G, int RPW, int NW, bool FMA, int PF = 1>
__global__ void __launch_bounds__(64 * NW)
stencil_kernel(const typename Traits<S>::T *__restrict__ src, typename Traits<S>::T *__restrict__ dst,
               Geometry geo, Coef<typename Traits<S>::C> k) {
  using R = TileRunner<S, VEC, T, G, RPW, NW, FMA, PF>;
  __shared__ typename Traits<S>::C xch[R::X::kElems];
  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  const int tx = lb % geo.ntx;
  const int rest = lb / geo.ntx;
  const int ty = rest % geo.nty;
  const int tz = rest / geo.nty;
  R tr(src, dst, geo, k, xch, tx, ty, tz);
  tr.run();
}

// The 2D 5-point operator on an nx x ny field streamed RPW rows per plane
// (geo.ny = RPW, geo.nz = planes, geo.rows2d = ny): one wave per workgroup.
template <int S, int VEC, int T, int G, int RPW, bool FMA, int PF = 1>
__global__ void __launch_bounds__(64)
stencil2d_kernel(const typename Traits<S>::T *__restrict__ src, typename Traits<S>::T *__restrict__ dst,
                 Geometry geo, Coef<typename Traits<S>::C> k) {
  using R = TileRunner<S, VEC, T, G, RPW, 1, FMA, PF, true>;
  using RR = TileRunner<S, VEC, T, G, RPW, 1, FMA, PF, true, true>;
  __shared__ typename Traits<S>::C xch[R::X::kElems];
  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  const int tx = lb % geo.ntx;
  const int tz = lb / geo.ntx;  // one row-block tile per x tile (nty = 1)
  // odd chunks stream backwards (see TileRunner)
  if (tz & 1) {
    RR tr(src, dst, geo, k, xch, tx, 0, tz);
    tr.run();
  } else {
    R tr(src, dst, geo, k, xch, tx, 0, tz);
    tr.run();
  }
}

// One member problem of a batched launch (device memory).  Members of one
// launch share the kernel (tile shape, dtype, depth); each has its own grid,
// buffers and coefficients.
struct BatchMember {
  void *buf[2];     // ping-pong buffers
  Geometry geo;     // launch geometry of this member
  double coef[7];   // ce cw cn cs ct cb cc
  int block_start;  // first logical block of this member in the launch
  int parity0;      // buffer holding the member's input at launch 0
};

// Batched launch: logical block -> member through a block table (one scalar
// load), then the same tile runner.  Launch `launch` of a sequence reads
// buf[(parity0 + launch) & 1] and writes the other buffer.
template <int S, int VEC, int T, int G, int RPW, int NW, bool FMA, int PF = 1>
__global__ void __launch_bounds__(64 * NW)
stencil_batch_kernel(const BatchMember *__restrict__ members, const int *__restrict__ block_member,
                     int launch) {
  using R = TileRunner<S, VEC, T, G, RPW, NW, FMA, PF>;
  using C = typename Traits<S>::C;
  using ST = typename Traits<S>::T;
  __shared__ C xch[R::X::kElems];
  const int lb = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
  const int m = block_member[lb];
  const BatchMember &d = members[m];
  const Geometry geo = d.geo;
  const int local = lb - d.block_start;
  const int par = (d.parity0 + launch) & 1;
  Geometry geo_l = geo;
  geo_l.step0 += launch * T;
  const Coef<C> k{(C)d.coef[0], (C)d.coef[1], (C)d.coef[2], (C)d.coef[3],
                  (C)d.coef[4], (C)d.coef[5], (C)d.coef[6]};
  const int tx = local % geo.ntx;
  const int rest = local / geo.ntx;
  const int ty = rest % geo.nty;
  const int tz = rest / geo.nty;
  R tr(static_cast<const ST *>(d.buf[par]), static_cast<ST *>(d.buf[par ^ 1]), geo_l, k, xch, tx, ty,
       tz);
  tr.run();
}

}  // namespace d3h
