// pbf_kernels.hip - hand-written CDNA4 (gfx950) kernels of one PBF step.
//
// Layout: particles live in slots sorted by (cell, particle id); each
// scene's slots start at a multiple of kSlotAlign.  A cell table
// cell_start (exclusive scan of per-cell counts) gives every cell's slot
// range, and because cell ids run x-fastest the x-cells of one (y, z) row of
// the stencil are ONE contiguous slot range: a search walks 9 ranges
// instead of 27 lists.  PAPER bins x in kXSub-times finer cells, so each
// range can be cut to the particle's reach.
//
// Step (reference sequence src/Simulation.cpp:455-576):
//   predict_count  predict (predict_positions.cl:5-9), cell id
//                  (update_cells.cl:9-15), per-cell arrival rank (atomic)
//   scan x3        exclusive scan of counts -> cell_start, counts zeroed
//   scatter        slot = cell_start + rank   (arrival order)
//   reorder        in-cell rank by particle id -> deterministic (cell, id)
//                  order; gathers x*, x, id into sorted slots
//   REF: I x (lambda, delta), each searching the 27-cell stencil
//                  (compute_scaling.cl, compute_delta.cl +
//                  update_predicted.cl); the last delta pass also forms
//                  v = (x* - x)/dt (update_velocities.cl); xsph_final: XSPH
//                  (apply_vorticity_and_viscosity.cl) fused with
//                  update_positions.cl.
//   PAPER: the first lambda pass searches and records the neighbour lists
//                  of the step; the other lambda and delta passes, XSPH +
//                  omega and the vorticity pass (fused with
//                  update_positions.cl) replay them.
#include "pbf_kernels.hpp"
#include "pbf_refmath.hpp"

#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

namespace pbf {

#define PBF_REF_EXACT _Pragma("clang fp contract(off)")

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Work-groups are dispatched round-robin over the 8 XCDs.  Remap so that
// each XCD walks one contiguous chunk of slots: spatially adjacent
// particles (which share stencil rows) then hit the same XCD's L2.  t is a
// position in dispatch order among nt tiles; t & 7 is its XCD, which gets
// the (t & 7)-th of 8 contiguous chunks of the nt tiles.
__device__ __forceinline__ int tile_of(int t, int nt) {
  const int q = nt >> 3, r = nt & 7, x = t & 7, k = t >> 3;
  return x * q + min(x, r) + k;
}

__device__ __forceinline__ int logical_block() {
  return tile_of(blockIdx.x, gridDim.x);
}

// The tiles of a launch (a tile: the blockDim.x slots a work-group
// covers), for (Tiles<PT> T(scenes); T.more(); T.next()).  PT false: one
// tile per work-group, gridDim.x tiles.  PT true (persistent launch):
// gridDim.x work-groups walk all the tiles of the handle's slots,
// work-group b the positions b, b + gridDim.x, ...; with gridDim.x a
// multiple of 8 those positions share b's XCD, so each work-group stays
// within its XCD's chunk.  Each position maps to one tile.
template <bool PT>
struct Tiles {
  int t, nt;
  __device__ __forceinline__ explicit Tiles(const Scene *__restrict__ scenes)
      : t(PT ? (int)blockIdx.x : 0),
        nt(PT ? scenes[0].slots_all / (int)blockDim.x : 1) {}
  __device__ __forceinline__ bool more() const { return t < nt; }
  __device__ __forceinline__ int tile() const {
    return PT ? tile_of(t, nt) : logical_block();
  }
  __device__ __forceinline__ void next() { t += PT ? (int)gridDim.x : 1; }
};

__device__ __forceinline__ int scene_of(const int *__restrict__ block_scene,
                                        int lb) {
  return __builtin_amdgcn_readfirstlane(
      block_scene[(lb * (int)blockDim.x) / kSlotAlign]);
}

// The kernels' "(int)(float)" made total: NaN -> 0, bounded magnitude.
__device__ __forceinline__ int trunc_int(float v) {
  if (v != v) return 0;
  v = fminf(fmaxf(v, -1e9f), 1e9f);
  return (int)v;
}

// 0 < r2 < h2 as one unsigned compare of the bit patterns: r2 is a sum of
// squares (+0 or positive), h2 > 0 is finite, and the order of positive
// floats is the order of their bits; +0 wraps to 0xffffffff, NaN and inf
// compare above h2.
__device__ __forceinline__ bool in_reach(float r2, float h2) {
  return __float_as_uint(r2) - 1u < __float_as_uint(h2) - 1u;
}

// r2 if in reach, else h2: there W = 0 exactly and h - r is within a
// rounding of 0 (exactly 0 where h2 is a power of 2), so an entry out of
// reach contributes nothing above the rounding of a sum.
__device__ __forceinline__ float in_reach_or_h2(float r2, float h2) {
  return in_reach(r2, h2) ? r2 : h2;
}

// (h - r)^2 / r, the spiky gradient's magnitude over r, from inv = 1/r:
// h^2/r - 2 h + r = (h^2 + r^2) inv - 2 h, one add and one fused
// multiply-add (r = 0 gives inf; callers select it away with in_reach).
__device__ __forceinline__ float spiky_over_r(float inv, float r2, float h,
                                              float h2) {
  return fmaf(inv, h2 + r2, -2.0f * h);
}

// Mixed precision records.  A position is 16-bit fixed point in units of
// q_unit = h / 2^kQBits from the scene's lower corner, each coordinate kept
// modulo 2^16, packed (x | y << 16, z | payload << 16); payloads and the
// velocities of the XSPH record are FP16.
__device__ __forceinline__ unsigned q16(float v, float lo, float scale) {
  return (unsigned)__float2int_rn((v - lo) * scale) & 0xffffu;
}

// FP16 bits of v, saturated at FP16's largest finite value (+-65504)
// rather than rounded to infinity: a near-coincident pair can give an
// |omega| or a velocity beyond FP16's range, which the FP32 passes carry as
// a large finite value; a NaN stays NaN (the state's own check reports it).
__device__ __forceinline__ unsigned f16_bits(float v) {
  const float c = v != v ? v : fminf(fmaxf(v, -65504.0f), 65504.0f);
  return (unsigned)__half_as_ushort(__float2half_rn(c));
}

__device__ __forceinline__ float f16_lo(unsigned w) {
  return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
}

__device__ __forceinline__ float f16_hi(unsigned w) {
  return __half2float(__ushort_as_half((unsigned short)(w >> 16)));
}

__device__ __forceinline__ uint2 pack_q(const Scene &sc, float3 x,
                                        unsigned payload) {
  const unsigned qx = q16(x.x, sc.bmin[0], sc.q_scale);
  const unsigned qy = q16(x.y, sc.bmin[1], sc.q_scale);
  const unsigned qz = q16(x.z, sc.bmin[2], sc.q_scale);
  return make_uint2(qx | (qy << 16), qz | (payload << 16));
}

// 1/rho of the XSPH weights, 0 for a particle with no density (no pair
// within h in the last lambda pass: it carries no weight).
__device__ __forceinline__ float inv_density(float rho) {
  return rho > 0.0f ? __builtin_amdgcn_rcpf(rho) : 0.0f;
}

// XSPH record: position, rho0 / rho (0 for no density; at most FP16's
// largest finite value, 65504) and v.
__device__ __forceinline__ uint4 pack_qv(const Scene &sc, float3 x, float rho,
                                         float3 v) {
  const uint2 q = pack_q(sc, x, f16_bits(sc.rho0 * inv_density(rho)));
  return make_uint4(q.x, q.y, f16_bits(v.x) | (f16_bits(v.y) << 16),
                    f16_bits(v.z));
}

// a - b of two packed positions in units of q_unit: each coordinate
// difference taken modulo 2^16 as a signed 16-bit value (exact for the
// quantised points while they are less than 2^15 units, 8 h, apart).
typedef short Short2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float3 q_diff(unsigned ax, unsigned ay,
                                         unsigned bx, unsigned by) {
  // x and y in one packed 16-bit subtraction (each half modulo 2^16)
  const Short2 dxy =
      __builtin_bit_cast(Short2, ax) - __builtin_bit_cast(Short2, bx);
  const int dz = (int)(short)(unsigned short)(ay - by);
  return make_float3((float)dxy.x, (float)dxy.y, (float)dz);
}

// x, y, z of a record: a float4 record read by a 12-byte load at the
// 16-byte stride (Gather<Pos3>), or a packed 12-byte record
// (Gather<Packed3>), for the passes that need only a neighbour's position.
struct Pos3 {
  float x, y, z;
};

// Positions packed at a 12-byte stride (x* for the search's candidates,
// x after a delta pass for the lambda replays): more records per cache
// line, so a wave's gather instruction touches fewer lines.
struct Packed3 {};

// A 12-byte record at 4-byte alignment (the stores of packed positions).
struct __attribute__((packed, aligned(4))) Float3P {
  float x, y, z;
};

typedef unsigned U3 __attribute__((ext_vector_type(3)));

// Gathers through a buffer resource: a 32-bit byte offset per lane instead
// of a 64-bit address, and loads past the n records of the array return 0.
template <typename T>
struct Gather {
  static constexpr bool kPos =
      std::is_same<T, Pos3>::value || std::is_same<T, Packed3>::value;
  // the value a record is read as
  using V = std::conditional_t<std::is_same<T, Packed3>::value, Pos3, T>;
  // bytes between records (Pos3 reads float4 records)
  static constexpr int kStride = std::is_same<T, Pos3>::value      ? 16
                                 : std::is_same<T, Packed3>::value ? 12
                                                                   : (int)sizeof(T);
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Gather(const void *p, int n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                            n * kStride, 0x00020000)) {}
  __device__ __forceinline__ V at_bytes(int off) const {
    if constexpr (kPos) {
      const U3 v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
      return Pos3{__uint_as_float(v.x), __uint_as_float(v.y),
                  __uint_as_float(v.z)};
    } else if constexpr (sizeof(T) == 16) {
      return __builtin_bit_cast(
          T, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    } else {
      static_assert(sizeof(T) == 8, "16- or 8-byte records");
      return __builtin_bit_cast(
          T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
  }
  __device__ __forceinline__ V operator[](int j) const {
    return at_bytes(j * kStride);
  }
  // records j .. j + 3, all their loads issued before any is used; packed
  // positions: 48 contiguous bytes, three 16-byte loads (4-byte aligned)
  __device__ __forceinline__ void batch4(int j, V &a0, V &a1, V &a2,
                                         V &a3) const {
    if constexpr (std::is_same<T, Packed3>::value) {
      const int o = j * 12;
      const uint4 u = __builtin_bit_cast(
          uint4, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0));
      const uint4 v = __builtin_bit_cast(
          uint4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 16, 0, 0));
      const uint4 w = __builtin_bit_cast(
          uint4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 32, 0, 0));
      a0 = Pos3{__uint_as_float(u.x), __uint_as_float(u.y),
                __uint_as_float(u.z)};
      a1 = Pos3{__uint_as_float(u.w), __uint_as_float(v.x),
                __uint_as_float(v.y)};
      a2 = Pos3{__uint_as_float(v.z), __uint_as_float(v.w),
                __uint_as_float(w.x)};
      a3 = Pos3{__uint_as_float(w.y), __uint_as_float(w.z),
                __uint_as_float(w.w)};
    } else {
      a0 = (*this)[j];
      a1 = (*this)[j + 1];
      a2 = (*this)[j + 2];
      a3 = (*this)[j + 3];
    }
  }
};

// XSPH's neighbour record from two arrays: x from the packed positions (12
// bytes per slot) and (v, 1/rho) (16 bytes per slot), two loads that touch
// fewer cache lines than the two halves of one 32-byte record would.
// Offsets are in units of 4 bytes per slot.
struct PosVel {};

template <>
struct Gather<PosVel> {
  using V = XVRec;
  static constexpr int kStride = 4;
  __amdgpu_buffer_rsrc_t rx, rv;
  __device__ __forceinline__ Gather(const float *x, const float4 *v, int n)
      : rx(__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(x), (short)0,
                                             n * 12, 0x00020000)),
        rv(__builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(v),
                                             (short)0, n * 16, 0x00020000)) {}
  __device__ __forceinline__ V at_bytes(int off) const {
    const U3 p = __builtin_amdgcn_raw_buffer_load_b96(rx, off * 3, 0, 0);
    const float4 q = __builtin_bit_cast(
        float4, __builtin_amdgcn_raw_buffer_load_b128(rv, off * 4, 0, 0));
    return XVRec{make_float4(__uint_as_float(p.x), __uint_as_float(p.y),
                             __uint_as_float(p.z), q.w),
                 make_float4(q.x, q.y, q.z, 0.0f)};
  }
  __device__ __forceinline__ V operator[](int j) const {
    return at_bytes(j * kStride);
  }
};

// A packed position by a plain load (a particle's own).
__device__ __forceinline__ float3 packed_at(const float *__restrict__ p,
                                            int i) {
  const Float3P v = reinterpret_cast<const Float3P *>(p)[i];
  return make_float3(v.x, v.y, v.z);
}

// Cell coordinates of a position.  REF: (int)((x - SYSTEM_MIN) /
// CELL_LENGTH) with IEEE division (update_cells.cl:9-15).  PAPER: the same
// with a multiply by the reciprocal.
template <int PH>
__device__ __forceinline__ void cell_coords(const Scene &sc, float x, float y,
                                            float z, int &cx, int &cy,
                                            int &cz) {
  if constexpr (PH == kRef) {
    cx = trunc_int((x - sc.bmin[0]) / sc.cell_len[0]);
    cy = trunc_int((y - sc.bmin[1]) / sc.cell_len[1]);
    cz = trunc_int((z - sc.bmin[2]) / sc.cell_len[2]);
  } else {
    cx = trunc_int((x - sc.bmin[0]) * sc.inv_cell[0]);
    cy = trunc_int((y - sc.bmin[1]) * sc.inv_cell[1]);
    cz = trunc_int((z - sc.bmin[2]) * sc.inv_cell[2]);
  }
}

// PAPER search rows.  The stencil is the 3 x 3 (y, z) cells around the
// particle's CURRENT cell (cells from its current position, lists from the
// step's sort, cells outside the grid skipped: compute_scaling.cl:21-53);
// each (y, z) cell is a row of fine x cells, contiguous in sorted order.  A
// row is cut to the fine x cells within the reach R = (1 + margin) h:
// with d the distance from xi to the row's y-z rectangle, the fine cells
// that meet [x - s, x + s], s = sqrt(R^2 - d^2), and no cells if d >= R.
// Every particle whose bin lies within R of xi is visited (the margin in R
// covers the rounding of this arithmetic).  Returns the sorted slot range
// [b, e) (b == e: nothing to visit).
__device__ __forceinline__ void row_range(const Scene &sc,
                                          const int *__restrict__ cs,
                                          float3 xi, int yy, int zz, int &b,
                                          int &e) {
  const float y0 = sc.bmin[1] + (float)yy * sc.cell_len[1];
  const float z0 = sc.bmin[2] + (float)zz * sc.cell_len[2];
  const float dy = fmaxf(fmaxf(y0 - xi.y, xi.y - (y0 + sc.cell_len[1])), 0.0f);
  const float dz = fmaxf(fmaxf(z0 - xi.z, xi.z - (z0 + sc.cell_len[2])), 0.0f);
  const float rem = sc.reach2 - dy * dy - dz * dz;
  b = e = 0;
  if (rem > 0.0f) {
    const float u = (xi.x - sc.bmin[0]) * sc.inv_cell[0];
    const float s = __builtin_amdgcn_sqrtf(rem) * sc.inv_cell[0];
    // clamp in float first: xi may lie far outside the grid
    const float lo = fmaxf(floorf(u - s), 0.0f);
    const float hi = fminf(floorf(u + s), (float)(sc.nc[0] - 1));
    if (lo <= hi) {
      const int base = sc.cell_base + (zz * sc.nc[1] + yy) * sc.nc[0];
      b = cs[base + (int)lo] + sc.shift;
      e = cs[base + (int)hi + 1] + sc.shift;
    }
  }
}

// Visit every slot of the 27-cell stencil of a particle at xi (cells from
// its CURRENT position, lists from the step's sort, neighbour cells outside
// the grid skipped: compute_scaling.cl:21-53).  REF: rows (dz, dy) each give
// one contiguous range over the x-cells [cx-1, cx+1] clipped to the grid.
// PAPER: the rows cut to the reach (row_range).
template <int PH, typename F>
__device__ __forceinline__ void for_stencil(const Scene &sc,
                                            const int *__restrict__ cs,
                                            float x, float y, float z, F &&f) {
  int cx, cy, cz;
  cell_coords<PH>(sc, x, y, z, cx, cy, cz);
  if constexpr (PH == kPaper) {
    const float3 xi = make_float3(x, y, z);
    for (int zz = max(cz - 1, 0); zz <= min(cz + 1, sc.nc[2] - 1); ++zz) {
      for (int yy = max(cy - 1, 0); yy <= min(cy + 1, sc.nc[1] - 1); ++yy) {
        int b, e;
        row_range(sc, cs, xi, yy, zz, b, e);
        for (int j = b; j < e; ++j) f(j);
      }
    }
    return;
  }
  const int xlo = max(cx - 1, 0), xhi = min(cx + 1, sc.nc[0] - 1);
  if (xlo > xhi) return;
  const int ylo = max(cy - 1, 0), yhi = min(cy + 1, sc.nc[1] - 1);
  const int zlo = max(cz - 1, 0), zhi = min(cz + 1, sc.nc[2] - 1);
  for (int zz = zlo; zz <= zhi; ++zz) {
    for (int yy = ylo; yy <= yhi; ++yy) {
      const int base = sc.cell_base + (zz * sc.nc[1] + yy) * sc.nc[0];
      const int b = cs[base + xlo] + sc.shift;
      const int e = cs[base + xhi + 1] + sc.shift;
      for (int j = b; j < e; ++j) f(j);
    }
  }
}

// REF traversal: the reference's exact visiting order, so that every sum is
// accumulated in the same order as the reference kernels: cells x-outer,
// z-inner (compute_scaling.cl:33-53), and within a cell descending particle
// id, the order of a list built by serial atomic_xchg insertion
// (update_cells.cl:18).  Slots within a cell are in ascending id, so each
// cell range is walked backwards.
template <typename F>
__device__ __forceinline__ void for_stencil_ref(const Scene &sc,
                                                const int *__restrict__ cs,
                                                float x, float y, float z,
                                                F &&f) {
  int cx, cy, cz;
  cell_coords<kRef>(sc, x, y, z, cx, cy, cz);
  for (int dx = -1; dx <= 1; ++dx) {
    const int nx = cx + dx;
    if (nx < 0 || nx >= sc.nc[0]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = cy + dy;
      if (ny < 0 || ny >= sc.nc[1]) continue;
      for (int dz = -1; dz <= 1; ++dz) {
        const int nz = cz + dz;
        if (nz < 0 || nz >= sc.nc[2]) continue;
        const int c = sc.cell_base + nx + (ny + nz * sc.nc[1]) * sc.nc[0];
        const int b = cs[c] + sc.shift;
        for (int j = cs[c + 1] + sc.shift - 1; j >= b; --j) f(j);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// REF pair terms, in the reference kernels' operation order, no contraction
// ---------------------------------------------------------------------------

// predictPositions (predict_positions.cl:8-9).
__device__ __forceinline__ float3 predict_exact(const Scene &sc, float4 x,
                                                float4 &v) {
  PBF_REF_EXACT
  v.x = v.x + sc.dt * 0.0f;
  v.y = v.y + sc.dt * sc.g_y;
  v.z = v.z + sc.dt * 0.0f;
  return make_float3(x.x + sc.dt * v.x, x.y + sc.dt * v.y, x.z + sc.dt * v.z);
}

// r^2 left to right without fusing (compute_scaling.cl:58-59).
__device__ __forceinline__ float dist2_exact(float3 xi, float4 xj, float &rx,
                                             float &ry, float &rz) {
  PBF_REF_EXACT
  rx = xi.x - xj.x;
  ry = xi.y - xj.y;
  rz = xi.z - xj.z;
  return rx * rx + ry * ry + rz * rz;
}

// compute_scaling.cl:56-84 for one list entry.
template <bool RN>
__device__ __forceinline__ bool ref_lambda_pair(const Scene &sc, float3 xi,
                                                float4 xj, float &dens,
                                                float &gsum, float &gix,
                                                float &giy, float &giz) {
  PBF_REF_EXACT
  float rx, ry, rz;
  const float r2 = dist2_exact(xi, xj, rx, ry, rz);
  const bool hit = r2 > 0.0f && r2 < sc.h2;
  if (hit) {
    float ux, uy, uz;
    const float rl =
        ref_unit<RN>(rx, ry, rz, r2, sc.h2 <= 0x1p40f, ux, uy, uz);
    const float gx = ux * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float gy = uy * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float gz = uz * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float w = sc.poly6 * (sc.h2 - r2) * (sc.h2 - r2) * (sc.h2 - r2);
    dens += w;
    gsum += ref_sqrt<RN>(gx * gx + gy * gy + gz * gz);
    gix += gx;
    giy += gy;
    giz += gz;
  }
  return hit;
}

// compute_scaling.cl:129-140.
__device__ __forceinline__ float ref_lambda_final(const Scene &sc, float dens,
                                                  float gsum, float gix,
                                                  float giy, float giz) {
  PBF_REF_EXACT
  gsum += sqrtf(gix * gix + giy * giy + giz * giz);
  const float c = (dens / sc.rho0) - 1.0f;
  return -1.0f * c / (gsum * gsum / (sc.rho0 * sc.rho0) + sc.eps);
}

// compute_delta.cl:56-80 for one list entry.
template <bool RN>
__device__ __forceinline__ void ref_delta_pair(const Scene &sc, float3 xi,
                                               float4 xj, float li, float lj,
                                               float &sx, float &sy,
                                               float &sz) {
  PBF_REF_EXACT
  float rx, ry, rz;
  const float r2 = dist2_exact(xi, xj, rx, ry, rz);
  if (r2 > 0.0f && r2 < sc.h2) {
    // -1 * rx / rl: -(rx / rl) (negation is exact)
    float ux, uy, uz;
    const float rl =
        ref_unit<RN>(rx, ry, rz, r2, sc.h2 <= 0x1p40f, ux, uy, uz);
    const float gx = -ux * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float gy = -uy * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float gz = -uz * sc.spiky * (sc.h - rl) * (sc.h - rl);
    const float s = li + lj;
    sx += s * gx;
    sy += s * gy;
    sz += s * gz;
  }
}

// compute_delta.cl:125-152 then update_predicted.cl:7: future = x* + dp,
// wall clamp (x-min wall moved by the wave generator), PAPER sphere first;
// returns x* + (future - x*).
__device__ __forceinline__ float3 clamp_update(const Scene &sc, float3 xi,
                                               float3 dp, bool sphere) {
  PBF_REF_EXACT
  float fx = xi.x + dp.x, fy = xi.y + dp.y, fz = xi.z + dp.z;
  if (sphere) {
    const float dx = fx - sc.sphere[0], dy = fy - sc.sphere[1],
                dz = fz - sc.sphere[2];
    const float rr = sc.sphere[3] + sc.h;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < rr * rr) {
      if (d2 > 0.0f) {
        const float s = rr / sqrtf(d2);
        fx = sc.sphere[0] + dx * s;
        fy = sc.sphere[1] + dy * s;
        fz = sc.sphere[2] + dz * s;
      } else {
        fy = sc.sphere[1] + rr;
      }
    }
  }
  if ((fx - sc.h) < (sc.bmin[0] + sc.wave)) {
    fx = sc.bmin[0] + sc.wave + sc.h;
  } else if ((fx + sc.h) > sc.bmax[0]) {
    fx = sc.bmax[0] - sc.h;
  }
  if ((fy - sc.h) < sc.bmin[1]) {
    fy = sc.bmin[1] + sc.h;
  } else if ((fy + sc.h) > sc.bmax[1]) {
    fy = sc.bmax[1] - sc.h;
  }
  if ((fz - sc.h) < sc.bmin[2]) {
    fz = sc.bmin[2] + sc.h;
  } else if ((fz + sc.h) > sc.bmax[2]) {
    fz = sc.bmax[2] - sc.h;
  }
  const float dx = fx - xi.x, dy = fy - xi.y, dz = fz - xi.z;
  return make_float3(xi.x + dx, xi.y + dy, xi.z + dz);
}

// apply_vorticity_and_viscosity.cl:56-61 for one list entry.
template <bool RN>
__device__ __forceinline__ void ref_xsph_pair(const Scene &sc, float3 xi,
                                              float4 xj, float4 vi, float4 vj,
                                              float &sx, float &sy,
                                              float &sz) {
  PBF_REF_EXACT
  float rx, ry, rz;
  const float r2 = dist2_exact(xi, xj, rx, ry, rz);
  if (r2 > 0.0f && r2 < sc.h2) {
    const float w = sc.poly6 * (sc.h2 - r2) * (sc.h2 - r2) * (sc.h2 - r2);
    const float inv = ref_recip<RN>(xj.w);
    sx += inv * (vj.x - vi.x) * w;
    sy += inv * (vj.y - vi.y) * w;
    sz += inv * (vj.z - vi.z) * w;
  }
}

// update_velocities.cl:8.
__device__ __forceinline__ float3 velocity_exact(const Scene &sc, float3 xs,
                                                 float4 xo) {
  PBF_REF_EXACT
  return make_float3((xs.x - xo.x) / sc.dt, (xs.y - xo.y) / sc.dt,
                     (xs.z - xo.z) / sc.dt);
}

// update_positions.cl:10-12: v += dv, |v|.
__device__ __forceinline__ float4 add_len_exact(float4 v, float dx, float dy,
                                                float dz, float &len) {
  PBF_REF_EXACT
  v.x = v.x + dx;
  v.y = v.y + dy;
  v.z = v.z + dz;
  len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return v;
}

__device__ __forceinline__ float4 mul_exact(float c, float x, float y,
                                            float z) {
  PBF_REF_EXACT
  return make_float4(c * x, c * y, c * z, 0.0f);
}

// ---------------------------------------------------------------------------
// neighbour search
// ---------------------------------------------------------------------------

// Slot i of the state (x, v): predict x* = x + dt (v + dt g) when
// `predict` (predict_positions.cl:5-9), bin it (update_cells.cl:9-15) and
// take an arrival rank in its cell.  Padding slots sit at the end of a
// scene's range and a wave never spans two scenes, so the valid lanes of a
// wave are a prefix of it; lanes past that prefix either pass valid = false
// or have left the kernel (the ballots below see them as invalid).
template <int PH>
__device__ __forceinline__ void predict_count_slot(
    const Scene &sc, int i, bool valid, float4 x, float4 v, int predict,
    float4 *__restrict__ pred_u, int *__restrict__ key,
    int *__restrict__ rank, int *__restrict__ count) {
  int k = -1;
  if (valid) {
    float3 p = make_float3(x.x, x.y, x.z);
    if (predict) p = predict_exact(sc, x, v);
    int cx, cy, cz;
    cell_coords<PH>(sc, p.x, p.y, p.z, cx, cy, cz);
    cx = min(max(cx, 0), sc.nc[0] - 1);
    cy = min(max(cy, 0), sc.nc[1] - 1);
    cz = min(max(cz, 0), sc.nc[2] - 1);
    k = sc.cell_base + cx + (cy + cz * sc.nc[1]) * sc.nc[0];
    key[i] = k;
    pred_u[i] = make_float4(p.x, p.y, p.z, v.w);
  }
  // Slots arrive in the previous step's cell order, so neighbouring lanes
  // mostly share a cell: one atomic per run of equal keys, the run's first
  // lane adding the run length and handing out ranks.
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(k, 1, 64);
  const bool head = valid && (lane == 0 || prev != k);
  const unsigned long long heads = __ballot(head);
  const unsigned long long stops = heads | ~__ballot(valid);
  const unsigned long long upto =
      lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int first = 63 - __clzll(heads & upto);  // valid lanes only
  const unsigned long long after = stops & ~upto;
  const int next = after ? __ffsll((long long)after) - 1 : 64;
  int base = 0;
  if (head) base = atomicAdd(&count[k], next - lane);
  base = __shfl(base, valid ? first : 0, 64);
  if (valid) rank[i] = base + lane - first;
}

template <int PH>
__global__ __launch_bounds__(256) void k_predict_count(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
    float4 *__restrict__ pred_u, int *__restrict__ key,
    int *__restrict__ rank, int *__restrict__ count, int predict) {
  const int lb = logical_block();
  const Scene &sc = scenes[scene_of(block_scene, lb)];
  const int i = lb * blockDim.x + threadIdx.x;
  // every lane stays for the cross-lane steps of predict_count_slot
  const bool valid = i - sc.slot_off < sc.n;
  float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v = x;
  if (valid) {
    x = pos[i];
    v = vel[i];
  }
  predict_count_slot<PH>(sc, i, valid, x, v, predict, pred_u, key, rank,
                         count);
}

// Exclusive scan over a block of W waves (lds: W ints); *total receives
// the block sum.
template <int W>
__device__ __forceinline__ int block_excl_scan(int v, int *lds, int &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int t = lds[k];
    before += (k < w) ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return before + x - v;
}

__device__ __forceinline__ int block_excl_scan256(int v, int *lds, int &total) {
  return block_excl_scan<4>(v, lds, total);
}

// A tile of the cell-count scan: 256 work-items x V int4 (4 V consecutive
// cells per work-item, 1024 V cells per work-group).
template <int V>
__device__ __forceinline__ int tile_counts(const int4 *__restrict__ count,
                                           int4 (&c)[V]) {
  const int4 *p = count + ((size_t)blockIdx.x * 256 + threadIdx.x) * V;
  int s = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    c[k] = p[k];
    s += c[k].x + c[k].y + c[k].z + c[k].w;
  }
  return s;
}

template <int V>
__global__ __launch_bounds__(256) void k_scan_reduce(
    const int4 *__restrict__ count, int *__restrict__ partial,
    int *__restrict__ dense) {
  __shared__ int lds[4];
  // the dense-cell list of this step's reorder starts empty (its last
  // reader, the previous step's k_reorder, has finished)
  if (blockIdx.x == 0 && threadIdx.x < 2) dense[threadIdx.x] = 0;
  int4 c[V];
  int tot;
  block_excl_scan256(tile_counts<V>(count, c), lds, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One work-group of kTopThreads: exclusive scan of the nb tile totals in
// place, kTopIpt consecutive totals per work-item per round.  partial is
// allocated to whole rounds (its padding stays 0: only entries below nb are
// written), so the loads need no bounds tests.
__global__ __launch_bounds__(kTopThreads) void k_scan_top(
    int *__restrict__ partial, int nb) {
  __shared__ int lds[kTopThreads / 64];
  static_assert(kTopIpt == 16, "four int4 per work-item");
  int4 *p4 = reinterpret_cast<int4 *>(partial) + threadIdx.x * 4;
  constexpr int kChunk4 = kTopChunk / 4;  // int4 per round
  // the next round's totals are loaded before this round is scanned
  int4 nxt[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nxt[k] = p4[k];
  int carry = 0;
  for (int base = 0; base < nb; base += kTopChunk, p4 += kChunk4) {
    int4 cur[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) cur[k] = nxt[k];
    if (base + kTopChunk < nb) {
#pragma unroll
      for (int k = 0; k < 4; ++k) nxt[k] = p4[kChunk4 + k];
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += cur[k].x + cur[k].y + cur[k].z + cur[k].w;
    int tot;
    int run = block_excl_scan<kTopThreads / 64>(s, lds, tot) + carry;
    const int first = base + threadIdx.x * kTopIpt;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int4 o;
      o.x = run;
      o.y = o.x + cur[k].x;
      o.z = o.y + cur[k].y;
      o.w = o.z + cur[k].z;
      run = o.w + cur[k].w;
      const int idx = first + 4 * k;
      if (idx + 3 < nb) {
        p4[k] = o;
      } else {
        const int vals[4] = {o.x, o.y, o.z, o.w};
        for (int q = 0; q < 4; ++q)
          if (idx + q < nb) partial[idx + q] = vals[q];
      }
    }
    carry += tot;
  }
}

// cell_start = exclusive scan of the counts, tile by tile from the tile's
// offset (k_scatter zeroes the counts of the occupied cells afterwards).
template <int V>
__device__ __forceinline__ void scan_down_tile(
    const int4 *__restrict__ count, int4 *__restrict__ cell_start, int offset,
    int *lds) {
  int4 c[V];
  int tot;
  int run =
      block_excl_scan256(tile_counts<V>(count, c), lds, tot) + offset;
  int4 *q = cell_start + ((size_t)blockIdx.x * 256 + threadIdx.x) * V;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    q[k] = make_int4(run, run + c[k].x, run + c[k].x + c[k].y,
                     run + c[k].x + c[k].y + c[k].z);
    run += c[k].x + c[k].y + c[k].z + c[k].w;
  }
}

// After k_scan_top: partial holds the exclusive scan of the tile totals.
template <int V>
__global__ __launch_bounds__(256) void k_scan_down(
    const int4 *__restrict__ count, const int *__restrict__ partial,
    int4 *__restrict__ cell_start) {
  __shared__ int lds[4];
  scan_down_tile<V>(count, cell_start, partial[blockIdx.x], lds);
}

// Straight after k_scan_reduce (up to kScanSumTiles tiles): partial holds
// the tile totals, and each tile sums those before it.
template <int V>
__global__ __launch_bounds__(256) void k_scan_down_sum(
    const int4 *__restrict__ count, const int *__restrict__ partial,
    int4 *__restrict__ cell_start) {
  __shared__ int lds[4];
  int before = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += 256) before += partial[k];
  int offset;
  block_excl_scan256(before, lds, offset);
  scan_down_tile<V>(count, cell_start, offset, lds);
}

// Also zeroes the count of every occupied cell for the next step's count
// (the scan has read them; cells without a particle hold 0 already), and
// lists each cell of more than kDenseCell slots in dense (count, cursor,
// then (first slot, end) pairs; the slot of arrival rank 0 lists it) for
// the bucketed rank of k_reorder.
template <bool PT>
__global__ __launch_bounds__(256) void k_scatter(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const int *__restrict__ key, const int *__restrict__ rank,
    const int *__restrict__ pid, const int *__restrict__ cell_start,
    int *__restrict__ count, int *__restrict__ perm, int *__restrict__ ppid,
    int *__restrict__ dense, int *__restrict__ err) {
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const int k = key[i];
    count[k] = 0;
    const int r = rank[i];
    const int slot = cell_start[k] + r + sc.shift;
    // counts that disagree with the keys would place a slot outside its
    // scene: reported (error bit 2), never written
    if ((unsigned)(slot - sc.slot_off) >= (unsigned)sc.n) {
      atomicOr(err, 2);
      continue;
    }
    perm[slot] = i;
    ppid[slot] = pid[i];
    if (r == 0) {
      const int hi =
          min(cell_start[k + 1] + sc.shift, sc.slot_off + sc.n);
      if (hi - slot > kDenseCell) {
        const int q = atomicAdd(&dense[0], 1);
        reinterpret_cast<int2 *>(dense + 2)[q] = make_int2(slot, hi);
      }
    }
  }
}

// The rank of each slot of a dense cell among the cell by particle id, in
// O(m) for ids spread over their range instead of the O(m^2) count of
// k_reorder's tiles: the cell's ids are counted into kDenseBuckets buckets
// of equal width over [min id, max id], the counts scanned to bucket
// starts, the slots placed bucket by bucket into scratch (order within a
// bucket arbitrary), and a slot's rank is its bucket's start plus the
// number of smaller ids in its bucket.  The work-groups of k_reorder take
// the cells k_scatter listed from a cursor once their tiles are done, until
// the list is exhausted; writes what the tiles write.  Every work-item of
// the work-group calls it (work-group barriers).
// One table of kDenseBuckets ints serves the three phases of a cell (the
// counts, scanned in place into the bucket starts, then advanced as the
// scatter's cursors, after which bucket b spans [cur[b - 1], cur[b])), so
// the LDS a work-group holds does not limit the reorder's occupancy.
struct DenseLds {
  int cur[kDenseBuckets];
  int item, min_id, max_id;
};

__device__ __forceinline__ void rank_dense_cells(
    const int *__restrict__ perm, const int *__restrict__ ppid,
    const float4 *__restrict__ pred_u, const float4 *__restrict__ pos,
    int *__restrict__ pid_s, float4 *__restrict__ xa, float *__restrict__ xp,
    float4 *__restrict__ xold, int *__restrict__ dense,
    int2 *__restrict__ scratch, DenseLds &L) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ncell = __builtin_amdgcn_readfirstlane(dense[0]);
  if (ncell == 0) return;  // the usual case: no dense cell this step
  const int2 *cells = reinterpret_cast<const int2 *>(dense + 2);
  for (;;) {
    // the cursor only grows: once it has passed the list, a work-group
    // leaves without an atomic (the late ones of a large launch)
    if (tid == 0)
      L.item = __hip_atomic_load(&dense[1], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) >= ncell
                   ? ncell
                   : atomicAdd(&dense[1], 1);
    __syncthreads();
    const int item = L.item;
    if (item >= ncell) return;  // uniform: every work-group reaches it
    for (int b = tid; b < kDenseBuckets; b += nt) L.cur[b] = 0;
    if (tid == 0) {
      L.min_id = INT_MAX;
      L.max_id = INT_MIN;
    }
    __syncthreads();
    const int lo = cells[item].x, hi = cells[item].y;
    int mn = INT_MAX, mx = INT_MIN;
    for (int t = lo + tid; t < hi; t += nt) {
      mn = min(mn, ppid[t]);
      mx = max(mx, ppid[t]);
    }
    atomicMin(&L.min_id, mn);
    atomicMax(&L.max_id, mx);
    __syncthreads();
    const int pmin = L.min_id;
    const int width = (L.max_id - pmin) / kDenseBuckets + 1;
    for (int t = lo + tid; t < hi; t += nt)
      atomicAdd(&L.cur[(ppid[t] - pmin) / width], 1);
    __syncthreads();
    if (tid == 0) {
      int run = 0;
      for (int b = 0; b < kDenseBuckets; ++b) {
        const int c = L.cur[b];
        L.cur[b] = run;
        run += c;
      }
    }
    __syncthreads();
    for (int t = lo + tid; t < hi; t += nt) {
      const int p = ppid[t];
      const int b = (p - pmin) / width;
      scratch[lo + atomicAdd(&L.cur[b], 1)] = make_int2(p, t);
    }
    __threadfence_block();
    __syncthreads();
    for (int q = lo + tid; q < hi; q += nt) {
      const int2 e = scratch[q];
      const int b = (e.x - pmin) / width;
      const int s0 = b > 0 ? L.cur[b - 1] : 0;
      const int b0 = lo + s0, b1 = lo + L.cur[b];
      int r = s0;
      for (int u = b0; u < b1; ++u) r += scratch[u].x < e.x ? 1 : 0;
      const int f = lo + r;
      const int src = perm[e.y];
      const float4 pu = pred_u[src];
      const float4 x = pos[src];
      pid_s[f] = e.x;
      if (xa) xa[f] = make_float4(pu.x, pu.y, pu.z, 0.0f);
      if (xp) reinterpret_cast<Float3P *>(xp)[f] = Float3P{pu.x, pu.y, pu.z};
      xold[f] = make_float4(x.x, x.y, x.z, pu.w);
    }
    __syncthreads();
  }
}

// Arrival order within a cell is not deterministic; rank each slot among
// its cell by particle id and move it there, gathering x*, x and the id.
// A cell of more than kDenseCell slots is skipped by the tiles and ranked
// by rank_dense_cells once a work-group's tiles are done.
template <bool PT>
__global__ __launch_bounds__(256) void k_reorder(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const int *__restrict__ perm, const int *__restrict__ ppid,
    const int *__restrict__ key, const int *__restrict__ cell_start,
    const float4 *__restrict__ pred_u, const float4 *__restrict__ pos,
    int *__restrict__ pid_s, float4 *__restrict__ xa, float *__restrict__ xp,
    float4 *__restrict__ xold, int *__restrict__ dense,
    int2 *__restrict__ dsc) {
  __shared__ DenseLds dl;
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int t = lb * blockDim.x + threadIdx.x;
    if (t - sc.slot_off >= sc.n) continue;
    const int src = perm[t];
    const int p = ppid[t];
    const int c = key[src];
    // clamped to the scene (only counts that disagree with the keys, which
    // k_scatter reports, reach past it)
    const int lo = max(cell_start[c] + sc.shift, sc.slot_off);
    const int hi = min(cell_start[c + 1] + sc.shift, sc.slot_off + sc.n);
    if (hi - lo > kDenseCell) continue;  // ranked by rank_dense_cells
    int r = 0;
    for (int u = lo; u < hi; ++u) r += (ppid[u] < p) ? 1 : 0;
    const int f = lo + r;
    const float4 pu = pred_u[src];
    const float4 x = pos[src];
    pid_s[f] = p;
    if (xa) xa[f] = make_float4(pu.x, pu.y, pu.z, 0.0f);
    if (xp) reinterpret_cast<Float3P *>(xp)[f] = Float3P{pu.x, pu.y, pu.z};
    xold[f] = make_float4(x.x, x.y, x.z, pu.w);
  }
  rank_dense_cells(perm, ppid, pred_u, pos, pid_s, xa, xp, xold, dense, dsc,
                   dl);
}

// ---------------------------------------------------------------------------
// neighbour lists (PAPER)
//
// PAPER follows Algorithm 1 of Macklin & Mueller: the neighbours of a
// particle are found once per step, at the predicted positions x*,
// N_i = {j != i : 0 < |x*_i - x*_j|^2 < h^2}, and every pass of the step
// (lambda and delta of each iteration, XSPH, vorticity) sums over N_i at its
// own positions, a pair contributing while 0 < r < h.  The first lambda pass
// searches and records N_i in visiting order; the other passes replay it.
//
// An entry is kEntryBits bits: stencil row (dz, dy) in the top 4 bits and
// the slot's offset from the row's first candidate slot in the low kOffBits
// (rows are those of x*_i; the search stores each particle's 9 row starts
// for the replaying passes).  kPerGroup entries are packed into a 16-byte
// group in registers (a 128-bit shift chain) and each full group is one
// store; group g of slot i is at nl[g * P + i], so lanes store and load
// adjacent groups.  A particle with more than kMaxHits neighbours (or a row
// of more than 2^kOffBits candidates) is marked with a count of
// kMaxHits + 1; the replaying passes then walk its rows again with the
// search's test at x*.
// ---------------------------------------------------------------------------

struct HitRecorder {
  unsigned w0, w1, w2, w3;
  unsigned room;  // entries the group in the registers takes after the
                  // next one (kPerGroup - 1 .. 0)
  int g;          // full groups stored
  __device__ __forceinline__ void push(unsigned e) {
    w0 = __builtin_amdgcn_alignbit(w1, w0, kEntryBits);
    w1 = __builtin_amdgcn_alignbit(w2, w1, kEntryBits);
    w2 = __builtin_amdgcn_alignbit(w3, w2, kEntryBits);
    w3 = __builtin_amdgcn_alignbit(e, w3, kEntryBits);
  }
  // Record an entry (its low kEntryBits bits).  No branch on the capacity:
  // past kMaxHits the groups overwrite the last one, which the overflow
  // mark (count kMaxHits + 1) makes unused.
  // The group is full when room, counted down once per entry, borrows: one
  // subtract whose borrow is the branch condition.
  __device__ __forceinline__ void add(unsigned entry, uint4 *__restrict__ nl,
                                     int stride, int i) {
    push(entry);
    if (__builtin_sub_overflow(room, 1u, &room)) {
      nl[(size_t)min(g, kMaxGroups - 1) * stride + i] =
          make_uint4(w0, w1, w2, w3);
      ++g;
      room = kPerGroup - 1;
    }
  }
  // Store the last partial group and the count.  The group is padded with
  // entries naming slot i itself (stencil row 4 is the particle's own row;
  // self_off = i - that row's first candidate slot), which every replaying
  // pass drops (r = 0), so a replay decodes all entries of a group without
  // a count test.  wide: a row has more candidates than an entry can
  // address, so the list is marked as overflowed.
  __device__ __forceinline__ void flush(uint4 *__restrict__ nl,
                                       int *__restrict__ ncnt, int stride,
                                       int i, int self_off, bool wide) {
    const int part = kPerGroup - 1 - (int)room;  // entries of the partial group
    const int c = g * kPerGroup + part;
    const bool over =
        wide || (part > 0 && (unsigned)self_off >= (1u << kOffBits));
    if (over || c > kMaxHits) {
      ncnt[i] = kMaxHits + 1;
      return;
    }
    ncnt[i] = c;
    if (part > 0) {
      const unsigned pad = (4u << kOffBits) | (unsigned)self_off;
      for (int k = part; k < kPerGroup; ++k) push(pad);
      nl[(size_t)g * stride + i] = make_uint4(w0, w1, w2, w3);
    }
  }
};

// Stencil walk of a PAPER searching pass over the rows of the 9 that lie in
// the grid: on_row(row, b, e) once per row ([b, e) its candidate slots, b
// the base of its list entries), then f(j, X[j], (row << kOffBits) - b)
// for every candidate, whose list entry is that value + j; candidate
// positions are loaded four at a time so that four loads are in flight per
// wait.
template <typename G, typename R, typename F>
__device__ __forceinline__ void walk_rows(const Scene &sc,
                                          const int *__restrict__ cs,
                                          const G &X, float3 xi,
                                          R &&on_row, F &&f) {
  int cx, cy, cz;
  cell_coords<kPaper>(sc, xi.x, xi.y, xi.z, cx, cy, cz);
  for (int dz = 0; dz < 3; ++dz) {
    const int zz = cz - 1 + dz;
    if (zz < 0 || zz >= sc.nc[2]) continue;
    for (int dy = 0; dy < 3; ++dy) {
      const int yy = cy - 1 + dy;
      if (yy < 0 || yy >= sc.nc[1]) continue;
      const int row = dz * 3 + dy;
      int b, e;
      row_range(sc, cs, xi, yy, zz, b, e);
      on_row(row, b, e);
      const int re = (row << kOffBits) - b;
      int j = b;
      for (; j + 4 <= e; j += 4) {
        Pos3 a0, a1, a2, a3;
        X.batch4(j, a0, a1, a2, a3);
        f(j, a0, re);
        f(j + 1, a1, re);
        f(j + 2, a2, re);
        f(j + 3, a3, re);
      }
      if (j < e) {
        // the 1 to 3 slots left, loaded together (a slot past e is loaded
        // but not visited; past the array a load returns 0) rather than
        // one load in flight at a time
        const Pos3 a0 = X[j], a1 = X[j + 1], a2 = X[j + 2];
        f(j, a0, re);
        if (j + 1 < e) f(j + 1, a1, re);
        if (j + 2 < e) f(j + 2, a2, re);
      }
    }
  }
}

// Entry q of a list group (bits kGroupSkip + q kEntryBits up; three of the
// eleven straddle two words).  q is a constant after unrolling.
__device__ __forceinline__ unsigned group_entry(const unsigned (&ws)[4],
                                                int q) {
  const int p = kGroupSkip + q * kEntryBits;
  const int wi = p >> 5, sh = p & 31;
  if (sh + kEntryBits <= 32) return ws[wi] >> sh;
  return __builtin_amdgcn_alignbit(ws[wi + 1], ws[wi], sh);
}

// f(X[j]) for every j of N_i in recorded order, the entries of a
// group resolved and all their loads issued before any is used (more loads
// in flight per wave measured faster than fewer VGPRs and more waves); the
// padding of the last group names i itself, which every pass's 0 < r^2
// test (or j != i) drops.  rbase holds the 9 row starts the search stored;
// rb is LDS scratch of 9 x blockDim.x ints.
template <typename T, typename F>
__device__ __forceinline__ void for_neighbours(
    const Scene &sc, const int *__restrict__ cs, const Gather<Packed3> &XS,
    const Gather<T> &X, const uint4 *__restrict__ nl,
    const int *__restrict__ ncnt, const int *__restrict__ rbase, int stride,
    int i, int *rb, F &&f) {
  constexpr int kS = Gather<T>::kStride;
  const int cnt = ncnt[i];
  if (cnt <= kMaxHits) {
    // rb[row * bs + lane] = the row's start in bytes of X (lanes of one row
    // in consecutive banks), from the starts the search stored
    const unsigned bs = blockDim.x;
    int *my = rb + threadIdx.x;
#pragma unroll
    for (int r = 0; r < 9; ++r)
      my[r * bs] = rbase[(size_t)r * stride + i] * kS;
    // table address row * 4 bs + 4 lane (one 24-bit multiply-add); record
    // byte offset start + entry offset * stride (one shift-add or 24-bit
    // multiply-add)
    const unsigned bs4 = bs * 4u, lane4 = threadIdx.x * 4u;
    const char *rbb = reinterpret_cast<const char *>(rb);
    constexpr unsigned kOffMask = (1u << kOffBits) - 1u;
    for (int g = 0; g * kPerGroup < cnt; ++g) {
      const uint4 w = nl[(size_t)g * stride + i];
      const unsigned ws[4] = {w.x, w.y, w.z, w.w};
      int off[kPerGroup];
#pragma unroll
      for (int q = 0; q < kPerGroup; ++q) {
        const unsigned e = group_entry(ws, q);
        const unsigned row = (e >> kOffBits) & 0xfu;
        off[q] = *reinterpret_cast<const int *>(rbb + row * bs4 + lane4) +
                 (int)(e & kOffMask) * kS;
      }
      typename Gather<T>::V xv[kPerGroup];
#pragma unroll
      for (int q = 0; q < kPerGroup; ++q) xv[q] = X.at_bytes(off[q]);
#pragma unroll
      for (int q = 0; q < kPerGroup; ++q) f(xv[q]);
    }
  } else {
    const float h2 = sc.h2;
    const Pos3 s4 = XS[i];
    const float3 xs = make_float3(s4.x, s4.y, s4.z);
    walk_rows(
        sc, cs, XS, xs, [](int, int, int) {},
        [&](int j, Pos3 sj, int) {
          const float rx = xs.x - sj.x, ry = xs.y - sj.y, rz = xs.z - sj.z;
          const float r2 = rx * rx + ry * ry + rz * rz;
          if (in_reach(r2, h2)) f(X[j]);
        });
  }
}

// ---------------------------------------------------------------------------
// constraint solve, REF (reference arithmetic and order; lambda array,
// x* alternating between two buffers)
// ---------------------------------------------------------------------------

// REF walk with the pair arithmetic taken out of the candidate loop.  The
// walk (for_stencil_ref's order) only tests j != i, 0 < r^2 < h^2 and puts
// the slots that pass in a per-lane queue in LDS (q: kRefQueue x blockDim.x
// ints); whenever a lane's queue is full, every lane runs pair(j) over its
// queued slots in order and empties it, and once more at the end.  A lane
// thus sums over the same pairs in the same order as with for_stencil_ref,
// while the wave runs the costly IEEE pair arithmetic about once per hit of
// its busiest lane instead of once per candidate slot.  pair(j) repeats the
// test (same operations, same result) inside the reference pair function.
// Used by the lambda pass, whose pair term (two square roots, three
// divisions) is the costliest: about 5 % off a c1 step.  The delta and XSPH
// passes measured slower with it (their terms cost less than re-gathering the
// queued neighbours), as did queues of 8 and 32 slots.
constexpr int kRefQueue = 16;

template <typename F>
__device__ __forceinline__ void for_stencil_ref_queued(
    const Scene &sc, const int *__restrict__ cs, const float4 *__restrict__ X,
    float3 xi, int i, int *q, F &&pair) {
  const unsigned bs = blockDim.x;
  int *my = q + threadIdx.x;
  int n = 0;
  auto drain = [&] {
    for (int k = 0; k < n; ++k) pair(my[k * bs]);
    n = 0;
  };
  const float h2 = sc.h2;
  for_stencil_ref(sc, cs, xi.x, xi.y, xi.z, [&](int j) {
    float rx, ry, rz;
    const float r2 = dist2_exact(xi, X[j], rx, ry, rz);
    if (j != i && r2 > 0.0f && r2 < h2) my[n++ * bs] = j;
    if (__ballot(n == kRefQueue)) drain();
  });
  drain();
}

// compute_scaling.cl:12-140.
template <bool PT, bool RN>
__global__ __launch_bounds__(256) void k_lambda_ref(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ X, const int *__restrict__ cs,
    float *__restrict__ lam, float *__restrict__ rho) {
  extern __shared__ int rq_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float4 x4 = X[i];
    const float3 xi = make_float3(x4.x, x4.y, x4.z);
    float dens = 0.0f, gsum = 0.0f, gix = 0.0f, giy = 0.0f, giz = 0.0f;
    for_stencil_ref_queued(sc, cs, X, xi, i, rq_lds, [&](int j) {
      ref_lambda_pair<RN>(sc, xi, X[j], dens, gsum, gix, giy, giz);
    });
    lam[i] = ref_lambda_final(sc, dens, gsum, gix, giy, giz);
    rho[i] = dens;
  }
}

// compute_delta.cl:13-152 + update_predicted.cl.  LAST also forms
// v = (x* - x)/dt (update_velocities.cl:8) and carries rho of the last
// lambda pass in .w for the XSPH pass.
template <bool PT, bool LAST, bool RN>
__global__ __launch_bounds__(256) void k_delta_ref(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ X, const int *__restrict__ cs,
    const float *__restrict__ lam, const float *__restrict__ rho,
    const float4 *__restrict__ xold, float4 *__restrict__ Xn,
    float4 *__restrict__ V) {
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float4 x4 = X[i];
    const float3 xi = make_float3(x4.x, x4.y, x4.z);
    const float li = lam[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for_stencil_ref(sc, cs, xi.x, xi.y, xi.z, [&](int j) {
      if (j != i) ref_delta_pair<RN>(sc, xi, X[j], li, lam[j], sx, sy, sz);
    });
    const float3 dp = make_float3(sx / sc.rho0, sy / sc.rho0, sz / sc.rho0);
    const float3 xn = clamp_update(sc, xi, dp, false);
    Xn[i] = make_float4(xn.x, xn.y, xn.z, LAST ? rho[i] : 0.0f);
    if constexpr (LAST) {
      const float4 xo = xold[i];
      const float3 v = velocity_exact(sc, xn, xo);
      V[i] = make_float4(v.x, v.y, v.z, xo.w);
    }
  }
}

// Where the last kernel of a step leaves the next step's predicted
// positions, cell keys and arrival ranks, and adds the cell counts
// (predict_count_slot on the state it has just written): the next step's
// sort then starts with the scan.
struct NextSort {
  float4 *pred_u;
  int *key;
  int *rank;
  int *count;
};

// Final write of a slot: x = x*, v, |v| (update_positions.cl:9-12), id.
__device__ __forceinline__ void write_final(int i, float3 xs, float4 v,
                                            float len, float4 *__restrict__ pos,
                                            float4 *__restrict__ vel,
                                            int *__restrict__ pid,
                                            const int *__restrict__ pid_s,
                                            int *__restrict__ err) {
  pos[i] = make_float4(xs.x, xs.y, xs.z, len);
  vel[i] = v;
  pid[i] = pid_s[i];
  // len = |v| is non-finite if any velocity component is
  if (!(__builtin_isfinite(xs.x) && __builtin_isfinite(xs.y) &&
        __builtin_isfinite(xs.z) && __builtin_isfinite(len)))
    atomicOr(err, 1);
}

// apply_vorticity_and_viscosity.cl:12-104 fused with update_positions.cl.
template <bool PT, bool RN>
__global__ __launch_bounds__(256) void k_xsph_final_ref(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ X, const int *__restrict__ cs,
    const float4 *__restrict__ V, const int *__restrict__ pid_s,
    float4 *__restrict__ pos, float4 *__restrict__ vel,
    int *__restrict__ pid, int *__restrict__ err, NextSort ns) {
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float4 x4 = X[i];
    const float3 xi = make_float3(x4.x, x4.y, x4.z);
    const float4 vi = V[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for_stencil_ref(sc, cs, xi.x, xi.y, xi.z, [&](int j) {
      if (j != i) ref_xsph_pair<RN>(sc, xi, X[j], vi, V[j], sx, sy, sz);
    });
    const float4 dv = mul_exact(sc.c_xsph, sx, sy, sz);
    float len;
    const float4 v = add_len_exact(vi, dv.x, dv.y, dv.z, len);
    write_final(i, xi, v, len, pos, vel, pid, pid_s, err);
    predict_count_slot<kRef>(sc, i, true, make_float4(xi.x, xi.y, xi.z, len), v,
                             1, ns.pred_u, ns.key, ns.rank, ns.count);
  }
}

// ---------------------------------------------------------------------------
// constraint solve, PAPER
//
// Buffers: XS = x* (sorted; it fixes the bins, rows and neighbour lists of
// the step and is not written after the sort), B = (x, lambda) written by a
// lambda pass so that the delta pass reads one float4 per neighbour, and
// XC = x after a delta pass (packed, 12 bytes per slot).  Iteration k's
// lambda pass reads XS (k = 0, the search) or XC; its delta pass reads B and
// writes XC.
// ---------------------------------------------------------------------------

// Sums of (h^2-r^2)^3, (h-r)^4 and (h-r)^2 r/|r| over the pairs of one
// particle, constants applied once: density and lambda (eqs. 1, 2, 8, 9,
// 11).
struct LambdaSums {
  float wsum = 0.0f, g2 = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
  __device__ __forceinline__ void add(float h, float h2, float rx, float ry,
                                      float rz, float r2) {
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float hr = h - r2 * inv;
    const float t = h2 - r2;
    const float hr2 = hr * hr;
    wsum += t * t * t;
    g2 += hr2 * hr2;
    const float gs = hr2 * inv;
    gx += rx * gs;
    gy += ry * gs;
    gz += rz * gs;
  }
  // add() of a replayed entry with no branch: an entry out of reach (r = 0
  // for the padding, or r >= h after the positions moved) is evaluated at
  // r^2 = h^2, where t = 0 exactly and h - r is within a rounding of 0
  // (exactly 0 in units of q_unit, where h^2 = 2^24): its terms vanish to
  // below the rounding of the sums, with one select instead of three.
  __device__ __forceinline__ void add_if(bool in, float h, float h2, float rx,
                                         float ry, float rz, float r2) {
    add(h, h2, rx, ry, rz, in ? r2 : h2);
  }
  __device__ __forceinline__ void store(const Scene &sc, int i, float3 xi,
                                        float4 *__restrict__ B,
                                        float *__restrict__ rho) const {
    const float dens = sc.poly6 * wsum;
    const float sp2 = sc.spiky * sc.spiky;
    const float denom = sp2 * (gx * gx + gy * gy + gz * gz + g2) *
                            (sc.inv_rho0 * sc.inv_rho0) +
                        sc.eps;
    B[i] = make_float4(xi.x, xi.y, xi.z, -(dens * sc.inv_rho0 - 1.0f) / denom);
    rho[i] = dens;
  }
  // store() for sums formed in units of q_unit (h = 2^kQBits, as the
  // mixed-precision passes form them): (h^2-r^2)^3 scales as q_unit^6, the
  // gradient terms (h-r)^4 and ((h-r)^2 r/|r|)^2 as q_unit^4.
  __device__ __forceinline__ void store_q(const Scene &sc, int i, float3 xi,
                                          float4 *__restrict__ B,
                                          float *__restrict__ rho) const {
    const float u2 = sc.q_unit * sc.q_unit, u4 = u2 * u2;
    const float dens = sc.poly6 * (u4 * u2) * wsum;
    const float sp2 = sc.spiky * sc.spiky;
    const float denom = sp2 * u4 * (gx * gx + gy * gy + gz * gz + g2) *
                            (sc.inv_rho0 * sc.inv_rho0) +
                        sc.eps;
    B[i] = make_float4(xi.x, xi.y, xi.z, -(dens * sc.inv_rho0 - 1.0f) / denom);
    rho[i] = dens;
  }
};

// Iteration 0: searches the rows of x*_i, records N_i, and forms lambda.
template <bool PT>
__global__ __launch_bounds__(256) void k_lambda_paper(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XP, const int *__restrict__ cs,
    float4 *__restrict__ B,
    float *__restrict__ rho, uint4 *__restrict__ nl, int *__restrict__ ncnt,
    int *__restrict__ rbase, int stride) {
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float3 xi = packed_at(XP, i);
    HitRecorder rec{0u, 0u, 0u, 0u, kPerGroup - 1, 0};
    LambdaSums ls;
    const float h = sc.h, h2 = sc.h2;
    const Gather<Packed3> xpg(XP, stride);
    // the first candidate slot of every stencil row in the grid, for the
    // replaying passes (row 4, the particle's own, also for the padding);
    // rows outside the grid name no entry and get 0
    int self_off = -1;
    unsigned in_grid = 0;
    bool wide = false;
    walk_rows(
        sc, cs, xpg, xi,
        [&](int row, int b, int e) {
          rbase[(size_t)row * stride + i] = b;
          in_grid |= 1u << row;
          if (row == 4) self_off = i - b;
          wide = wide || e - b > (1 << kOffBits);
        },
        [&](int j, Pos3 xj, int re) {
          const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
          const float r2 = rx * rx + ry * ry + rz * rz;
          if (in_reach(r2, h2)) {
            rec.add((unsigned)(re + j), nl, stride, i);
            ls.add(h, h2, rx, ry, rz, r2);
          }
        });
  #pragma unroll
    for (int r = 0; r < 9; ++r)
      if (!(in_grid >> r & 1u)) rbase[(size_t)r * stride + i] = 0;
    rec.flush(nl, ncnt, stride, i, self_off, wide);
    ls.store(sc, i, xi, B, rho);
  }
}

// Iterations >= 1: lambda over N_i at the positions of the last delta pass.
template <bool PT>
__global__ __launch_bounds__(256) void k_lambda_list(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const int *__restrict__ cs, const uint4 *__restrict__ nl,
    const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, float4 *__restrict__ B,
    float *__restrict__ rho) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float3 xi = packed_at(XC, i);
    LambdaSums ls;
    const float h = sc.h, h2 = sc.h2;
    const Gather<Packed3> xsg(XS, stride);
    const Gather<Packed3> xcg(XC, stride);
    for_neighbours(sc, cs, xsg, xcg, nl, ncnt, rbase, stride, i, rb_lds,
                   [&](Pos3 xj) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = rx * rx + ry * ry + rz * rz;
      ls.add_if(in_reach(r2, h2), h, h2, rx, ry, rz, r2);
    });
    ls.store(sc, i, xi, B, rho);
  }
}

// The support radius and its square in units of q_unit (exact).
constexpr float kQH = (float)(1 << kQBits);
constexpr float kQH2 = kQH * kQH;

// Mixed precision, iterations 1 .. m - 1: k_lambda_list over 16-bit
// fixed-point neighbour positions (QA, written by the delta pass before),
// the sums formed in units of q_unit.
template <bool PT>
__global__ __launch_bounds__(256) void k_lambda_list_q(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const uint2 *__restrict__ QA, const int *__restrict__ cs,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, float4 *__restrict__ B,
    float *__restrict__ rho) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float3 xi = packed_at(XC, i);
    // the own record, quantised as the neighbours' are (the padding entries,
    // which name i itself, then give r = 0 exactly)
    const uint2 qi = QA[i];
    LambdaSums ls;
    const Gather<Packed3> xsg(XS, stride);
    const Gather<uint2> qag(QA, stride);
    for_neighbours(sc, cs, xsg, qag, nl, ncnt, rbase, stride, i, rb_lds,
                   [&](uint2 qj) {
      const float3 r = q_diff(qi.x, qi.y, qj.x, qj.y);
      const float r2 = r.x * r.x + r.y * r.y + r.z * r.z;
      ls.add_if(in_reach(r2, kQH2), kQH, kQH2, r.x, r.y, r.z, r2);
    });
    ls.store_q(sc, i, xi, B, rho);
  }
}

// eqs. 13, 14 and the collision clamp over N_i; writes XC.  LAST also
// writes VR = (v, 1/rho), v = (x - x_n)/dt, for the XSPH pass.
// WRITE_QA / WRITE_QV (mixed precision) also write the 16-bit records of
// the next lambda pass / of the XSPH pass.
template <bool PT, bool LAST, bool WRITE_QA, bool N4, bool WRITE_QV>
__global__ __launch_bounds__(256) void k_delta_paper(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float4 *__restrict__ B,
    const int *__restrict__ cs, const float *__restrict__ rho,
    const float4 *__restrict__ xold, const uint4 *__restrict__ nl,
    const int *__restrict__ ncnt, const int *__restrict__ rbase, int stride,
    float *__restrict__ XC, float4 *__restrict__ VR, uint2 *__restrict__ QA,
    uint4 *__restrict__ QV) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float4 x4 = B[i];
    const float3 xi = make_float3(x4.x, x4.y, x4.z);
    const float li = x4.w;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    const float h = sc.h, h2 = sc.h2, cq = sc.scorr_cq, kk = sc.scorr_k;
    const int n = sc.scorr_n;
    const Gather<Packed3> xsg(XS, stride);
    const Gather<float4> bg(B, stride);
    // (lambda_i + lambda_j + s_corr) (h-r)^2 r/|r| summed branch-free (the
    // out-of-reach entries selected to 0).  N4: the paper's exponent n = 4 in
    // every scene (the host picks the instance; the loop over any n costs
    // registers and occupancy)
    for_neighbours(sc, cs, xsg, bg, nl, ncnt, rbase, stride, i, rb_lds,
                   [&](float4 xj) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = rx * rx + ry * ry + rz * rz;
      const float inv = __builtin_amdgcn_rsqf(r2);
      const float t = h2 - r2;
      const float q = t * t * t * cq;  // W / W(dq)
      float qn;
      if constexpr (N4) {
        const float q2 = q * q;
        qn = q2 * q2;
      } else {
        qn = 1.0f;
        for (int e = 0; e < n; ++e) qn *= q;
      }
      const float coef = li + xj.w - kk * qn;
      const float gs =
          in_reach(r2, h2) ? coef * spiky_over_r(inv, r2, h, h2) : 0.0f;
      ax += rx * gs;
      ay += ry * gs;
      az += rz * gs;
    });
    // grad_i W = -spiky (h-r)^2 r/|r|
    const float c = -sc.spiky * sc.inv_rho0;
    const float3 dp = make_float3(ax * c, ay * c, az * c);
    const float3 xn = clamp_update(sc, xi, dp, sc.sphere[3] > 0.0f);
    if constexpr (WRITE_QA) QA[i] = pack_q(sc, xn, 0u);
    reinterpret_cast<Float3P *>(XC)[i] = Float3P{xn.x, xn.y, xn.z};
    if constexpr (LAST) {
      const float3 v = velocity_exact(sc, xn, xold[i]);
      VR[i] = make_float4(v.x, v.y, v.z, inv_density(rho[i]));
      if constexpr (WRITE_QV) QV[i] = pack_qv(sc, xn, rho[i], v);
    }
  }
}

// The XSPH sum over N_i: (v_j - v_i) W_ij / rho_j, a neighbour with no
// density (no pair within h in the last lambda pass) carrying no weight
// (its record holds 1/rho_j = 0); with OMEGA also omega_i = sum (v_j - v_i)
// x grad_j W_ij (eq. 15), spiky applied by the caller.  QIN: neighbours
// from the 16-bit records QV, the sums formed in units of q_unit (scale: s
// by q_unit^6 / rho0, since the records hold rho0 / rho; om by q_unit^2).
template <bool QIN, bool OMEGA>
__device__ __forceinline__ void xsph_sums(
    const Scene &sc, const int *__restrict__ cs, const Gather<Packed3> &xsg,
    const float *__restrict__ XC, const float4 *__restrict__ VR,
    const uint4 *__restrict__ QV,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, int i, int *rb, float3 xi,
    float4 vi, float3 &s, float3 &om) {
  const float h = QIN ? kQH : sc.h, h2 = QIN ? kQH2 : sc.h2;
  // inv_j: 1/rho_j (FP32 records) or rho0/rho_j (16-bit records), 0 for a
  // neighbour with no density
  auto term = [&](float rx, float ry, float rz, float inv_j, float3 vj) {
    // branch-free: an out-of-reach entry is evaluated at r^2 = h^2, where
    // W = 0 exactly and the spiky factor is within a rounding of 0
    const float r2 = in_reach_or_h2(rx * rx + ry * ry + rz * rz, h2);
    const float vx = vj.x - vi.x, vy = vj.y - vi.y, vz = vj.z - vi.z;
    const float t = h2 - r2;
    const float w = t * t * t * inv_j;
    s.x += vx * w;
    s.y += vy * w;
    s.z += vz * w;
    if constexpr (OMEGA) {
      const float inv = __builtin_amdgcn_rsqf(r2);
      // grad_j W_ij = spiky gs r
      const float gs = spiky_over_r(inv, r2, h, h2);
      const float gx = rx * gs, gy = ry * gs, gz = rz * gs;
      om.x += vy * gz - vz * gy;
      om.y += vz * gx - vx * gz;
      om.z += vx * gy - vy * gx;
    }
  };
  if constexpr (QIN) {
    const uint4 qi = QV[i];
    const Gather<uint4> qvg(QV, stride);
    for_neighbours(sc, cs, xsg, qvg, nl, ncnt, rbase, stride, i, rb,
                   [&](uint4 qj) {
      const float3 r = q_diff(qi.x, qi.y, qj.x, qj.y);
      term(r.x, r.y, r.z, f16_hi(qj.y),
           make_float3(f16_lo(qj.z), f16_hi(qj.z), f16_lo(qj.w)));
    });
  } else {
    const Gather<PosVel> xvg(XC, VR, stride);
    for_neighbours(sc, cs, xsg, xvg, nl, ncnt, rbase, stride, i, rb,
                   [&](XVRec nb) {
      term(xi.x - nb.x.x, xi.y - nb.x.y, xi.z - nb.x.z, nb.x.w,
           make_float3(nb.v.x, nb.v.y, nb.v.z));
    });
  }
}

// Factor of the XSPH sum: c poly6 (FP32 sums), c poly6 q_unit^6 / rho0
// (QIN sums).
template <bool QIN>
__device__ __forceinline__ float xsph_factor(const Scene &sc) {
  if constexpr (QIN) {
    const float u2 = sc.q_unit * sc.q_unit;
    return sc.c_xsph * sc.poly6 * (u2 * u2 * u2) * sc.inv_rho0;
  } else {
    return sc.c_xsph * sc.poly6;
  }
}

// XSPH over N_i with 1/rho_j weights fused with update_positions.cl (PAPER
// without vorticity).
template <bool PT, bool QIN>
__global__ __launch_bounds__(256) void k_xsph_final_paper(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const float4 *__restrict__ VR, const float4 *__restrict__ xold,
    const uint4 *__restrict__ QV, const int *__restrict__ cs,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride,
    const int *__restrict__ pid_s, float4 *__restrict__ pos,
    float4 *__restrict__ vel, int *__restrict__ pid, int *__restrict__ err,
    NextSort ns) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float3 xi = packed_at(XC, i);
    // own v and m (the mass from x_n's record)
    const float4 vr = VR[i];
    const float4 vi = make_float4(vr.x, vr.y, vr.z, xold[i].w);
    float3 s = make_float3(0.0f, 0.0f, 0.0f), om = s;
    const Gather<Packed3> xsg(XS, stride);
    xsph_sums<QIN, false>(sc, cs, xsg, XC, VR, QV, nl, ncnt, rbase, stride, i,
                          rb_lds, xi, vi, s, om);
    const float c = xsph_factor<QIN>(sc);
    const float4 v =
        make_float4(vi.x + c * s.x, vi.y + c * s.y, vi.z + c * s.z, vi.w);
    const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    write_final(i, xi, v, len, pos, vel, pid, pid_s, err);
    predict_count_slot<kPaper>(sc, i, true, make_float4(xi.x, xi.y, xi.z, len),
                               v, 1, ns.pred_u, ns.key, ns.rank, ns.count);
  }
}

// |omega| as the FP16 payload of the vorticity pass's records (scaled into
// FP16's range).
constexpr float kOmegaScale = 1.0f / 256.0f;

// XSPH and vorticity omega_i = sum (v_j - v_i) x grad_j W_ij (eq. 15) over
// N_i; writes B = (x, |omega|) for the vorticity pass, or (QIN) its 16-bit
// records QA.
template <bool PT, bool QIN>
__device__ __forceinline__ void xsph_omega_body(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const float4 *__restrict__ VR, const float4 *__restrict__ xold,
    const uint4 *__restrict__ QV, const int *__restrict__ cs,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, float4 *__restrict__ VX,
    float4 *__restrict__ OM, float4 *__restrict__ B,
    uint2 *__restrict__ QA) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    const float3 xi = packed_at(XC, i);
    // own v and m (the mass from x_n's record)
    const float4 vr = VR[i];
    const float4 vi = make_float4(vr.x, vr.y, vr.z, xold[i].w);
    float3 s = make_float3(0.0f, 0.0f, 0.0f), om = s;
    const Gather<Packed3> xsg(XS, stride);
    xsph_sums<QIN, true>(sc, cs, xsg, XC, VR, QV, nl, ncnt, rbase, stride, i,
                         rb_lds, xi, vi, s, om);
    const float cx = xsph_factor<QIN>(sc);
    VX[i] =
        make_float4(vi.x + cx * s.x, vi.y + cx * s.y, vi.z + cx * s.z, vi.w);
    const float cw = QIN ? sc.spiky * (sc.q_unit * sc.q_unit) : sc.spiky;
    const float ox = om.x * cw, oy = om.y * cw, oz = om.z * cw;
    const float wl = sqrtf(ox * ox + oy * oy + oz * oz);
    OM[i] = make_float4(ox, oy, oz, wl);
    if constexpr (QIN)
      QA[i] = pack_q(sc, xi, f16_bits(wl * kOmegaScale));
    else
      B[i] = make_float4(xi.x, xi.y, xi.z, wl);
  }
}

// The FP32 instance at the compiler's register count (4 waves per SIMD; at 5
// it spills); the 16-bit instance capped at 5 waves, which it fits without
// spilling (c4 mixed: XSPH + omega 283.5 against 294.9 us).
template <bool PT>
__global__ __launch_bounds__(256) void k_xsph_omega(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const float4 *__restrict__ VR, const float4 *__restrict__ xold,
    const uint4 *__restrict__ QV, const int *__restrict__ cs,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, float4 *__restrict__ VX,
    float4 *__restrict__ OM, float4 *__restrict__ B,
    uint2 *__restrict__ QA) {
  xsph_omega_body<PT, false>(scenes, block_scene, XS, XC, VR, xold, QV, cs, nl, ncnt, rbase, stride, VX, OM, B, QA);
}
template <bool PT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void
k_xsph_omega_q(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float *__restrict__ XC,
    const float4 *__restrict__ VR, const float4 *__restrict__ xold,
    const uint4 *__restrict__ QV, const int *__restrict__ cs,
    const uint4 *__restrict__ nl, const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, float4 *__restrict__ VX,
    float4 *__restrict__ OM, float4 *__restrict__ B,
    uint2 *__restrict__ QA) {
  xsph_omega_body<PT, true>(scenes, block_scene, XS, XC, VR, xold, QV, cs, nl, ncnt, rbase, stride, VX, OM, B, QA);
}

// eta = sum (|w_j| - |w_i|) grad_i W_ij over N_i, N = eta/|eta|,
// v += dt eps (N x w_i) (eq. 16), then update_positions.cl; reads
// B = (x, |omega|), or (QIN) the 16-bit records QA (eta formed in units of
// q_unit and of the FP16 scale: only its direction is used) and the own
// position from XC.
template <bool PT, bool QIN>
__global__ __launch_bounds__(256) void k_vort_final(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float *__restrict__ XS, const float4 *__restrict__ B,
    const float *__restrict__ XC, const uint2 *__restrict__ QA,
    const int *__restrict__ cs, const float4 *__restrict__ VX,
    const float4 *__restrict__ OM, const uint4 *__restrict__ nl,
    const int *__restrict__ ncnt,
    const int *__restrict__ rbase, int stride, const int *__restrict__ pid_s,
    float4 *__restrict__ pos, float4 *__restrict__ vel, int *__restrict__ pid,
    int *__restrict__ err, NextSort ns) {
  extern __shared__ int rb_lds[];
  for (Tiles<PT> T(scenes); T.more(); T.next()) {
    const int lb = T.tile();
    const Scene &sc = scenes[scene_of(block_scene, lb)];
    const int i = lb * blockDim.x + threadIdx.x;
    if (i - sc.slot_off >= sc.n) continue;
    float3 xi;
    if constexpr (QIN) {
      xi = packed_at(XC, i);
    } else {
      const float4 x4 = B[i];
      xi = make_float3(x4.x, x4.y, x4.z);
    }
    const float4 wi = OM[i];
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    const float h = QIN ? kQH : sc.h, h2 = QIN ? kQH2 : sc.h2;
    auto term = [&](float rx, float ry, float rz, float dw) {
      const float r2 = rx * rx + ry * ry + rz * rz;
      const float inv = __builtin_amdgcn_rsqf(r2);
      const float gs =
          in_reach(r2, h2) ? dw * spiky_over_r(inv, r2, h, h2) : 0.0f;
      ex += rx * gs;
      ey += ry * gs;
      ez += rz * gs;
    };
    const Gather<Packed3> xsg(XS, stride);
    if constexpr (QIN) {
      // |w_i| rounded as the neighbours' are, so that equal magnitudes cancel
      const uint2 qi = QA[i];
      const float wq = f16_hi(qi.y);
      const Gather<uint2> qag(QA, stride);
      for_neighbours(sc, cs, xsg, qag, nl, ncnt, rbase, stride, i, rb_lds,
                     [&](uint2 qj) {
        const float3 r = q_diff(qi.x, qi.y, qj.x, qj.y);
        term(r.x, r.y, r.z, f16_hi(qj.y) - wq);
      });
    } else {
      const Gather<float4> bg(B, stride);
      for_neighbours(sc, cs, xsg, bg, nl, ncnt, rbase, stride, i, rb_lds,
                     [&](float4 xj) {
        term(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z, xj.w - wi.w);
      });
    }
    float4 v = VX[i];
    const float e2 = ex * ex + ey * ey + ez * ez;
    if (e2 > 0.0f) {
      // the sign of grad_i W (-spiky) flips eta; N = -e/|e|
      const float inv = -__builtin_amdgcn_rsqf(e2);
      const float nx = ex * inv, ny = ey * inv, nz = ez * inv;
      const float c = sc.dt * sc.vort_eps;
      v.x += c * (ny * wi.z - nz * wi.y);
      v.y += c * (nz * wi.x - nx * wi.z);
      v.z += c * (nx * wi.y - ny * wi.x);
    }
    const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    write_final(i, xi, v, len, pos, vel, pid, pid_s, err);
    predict_count_slot<kPaper>(sc, i, true, make_float4(xi.x, xi.y, xi.z, len),
                               v, 1, ns.pred_u, ns.key, ns.rank, ns.count);
  }
}

// ---------------------------------------------------------------------------
// probes and diagnostics
// ---------------------------------------------------------------------------

// PAPER probe: lambda from B = (x*, lambda) into the lambda array.
__global__ __launch_bounds__(256) void k_unpack_lambda(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ B, float *__restrict__ lam) {
  const int lb = logical_block();
  const Scene &sc = scenes[scene_of(block_scene, lb)];
  const int i = lb * blockDim.x + threadIdx.x;
  if (i - sc.slot_off >= sc.n) return;
  lam[i] = B[i].w;
}


// Stencil neighbour count with the reference predicate 0 < r^2 < h^2,
// r^2 unfused (compute_scaling.cl:56-62).
template <int PH>
__global__ __launch_bounds__(256) void k_count_nbrs(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ X, const int *__restrict__ cs,
    int *__restrict__ nbr) {
  const int lb = logical_block();
  const Scene &sc = scenes[scene_of(block_scene, lb)];
  const int i = lb * blockDim.x + threadIdx.x;
  if (i - sc.slot_off >= sc.n) return;
  const float4 x4 = X[i];
  const float3 xi = make_float3(x4.x, x4.y, x4.z);
  int c = 0;
  for_stencil<PH>(sc, cs, xi.x, xi.y, xi.z, [&](int j) {
    float rx, ry, rz;
    const float r2 = dist2_exact(xi, X[j], rx, ry, rz);
    if (j != i && r2 > 0.0f && r2 < sc.h2) ++c;
  });
  nbr[i] = c;
}

// Density at the sorted positions X, then per-block FP64 partials, e =
// rho/rho0 - 1: [max e+ (all particles), sum |e| and count (particles with
// at least one neighbour), px, py, pz, energy, max |e| (particles with a
// neighbour)].
template <int PH>
__global__ __launch_bounds__(256) void k_diag(
    const Scene *__restrict__ scenes, const int *__restrict__ block_scene,
    const float4 *__restrict__ X, const int *__restrict__ cs,
    const float4 *__restrict__ vel, double *__restrict__ red) {
  __shared__ double lds[8][4];
  const int lb = logical_block();
  const Scene &sc = scenes[scene_of(block_scene, lb)];
  const int i = lb * blockDim.x + threadIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i - sc.slot_off < sc.n) {
    const float4 x4 = X[i];
    const float3 xi = make_float3(x4.x, x4.y, x4.z);
    float wsum = 0.0f;
    int pairs = 0;
    for_stencil<PH>(sc, cs, xi.x, xi.y, xi.z, [&](int j) {
      const float4 xj = X[j];
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = rx * rx + ry * ry + rz * rz;
      if (j != i && r2 > 0.0f && r2 < sc.h2) {
        const float t = sc.h2 - r2;
        wsum += t * t * t;
        ++pairs;
      }
    });
    const double e = (double)(sc.poly6 * wsum) / (double)sc.rho0 - 1.0;
    const float4 vv = vel[i];
    const double m = vv.w;
    v[0] = e > 0.0 ? e : 0.0;
    v[1] = pairs ? fabs(e) : 0.0;
    v[2] = pairs ? 1.0 : 0.0;
    v[3] = m * vv.x;
    v[4] = m * vv.y;
    v[5] = m * vv.z;
    v[6] = 0.5 * m * ((double)vv.x * vv.x + (double)vv.y * vv.y +
                      (double)vv.z * vv.z);
    v[7] = v[1];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = (k == 0 || k == 7) ? fmax(x, y) : x + y;
    }
    v[k] = x;
  }
  if (lane == 0)
    for (int k = 0; k < 8; ++k) lds[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    const int nw = blockDim.x >> 6;
    double x = lds[k][0];
    for (int q = 1; q < nw; ++q)
      x = (k == 0 || k == 7) ? fmax(x, lds[k][q]) : x + lds[k][q];
    red[(size_t)lb * 8 + k] = x;
  }
}

__global__ __launch_bounds__(256) void k_diag_final(double *__restrict__ red,
                                                    int nblocks,
                                                    double *__restrict__ out) {
  __shared__ double lds[8][4];
  double v[8];
  for (int k = 0; k < 8; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int k = 0; k < 8; ++k) {
      const double x = red[(size_t)b * 8 + k];
      v[k] = (k == 0 || k == 7) ? fmax(v[k], x) : v[k] + x;
    }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 8; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = (k == 0 || k == 7) ? fmax(x, y) : x + y;
    }
    v[k] = x;
  }
  if (lane == 0)
    for (int k = 0; k < 8; ++k) lds[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    double x = lds[k][0];
    for (int q = 1; q < 4; ++q)
      x = (k == 0 || k == 7) ? fmax(x, lds[k][q]) : x + lds[k][q];
    out[k] = x;
  }
}

// ---------------------------------------------------------------------------
// host-side launch sequencing
// ---------------------------------------------------------------------------

namespace {

struct Launcher {
  const StepShape &sh;
  hipStream_t st;
  KernelTimer *timer;
};

// Work-groups of block threads (smem bytes of LDS each) the device holds at
// once for kernel k, a multiple of 8 (one chunk per XCD); 0 if unknown.
int resident_grid(const void *k, int block, size_t smem, int cus) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, smem) !=
          hipSuccess ||
      per_cu <= 0)
    return 0;
  return per_cu * cus / 8 * 8;
}

}  // namespace

// With a timer, the kernel is launched with start/stop events attached to
// its dispatch (hipExtLaunchKernelGGL), which time the kernel itself rather
// than the gaps between launches.
#define PBF_LAUNCH_SMEM(L, cls, kern, grid, block, smem, ...)                \
  do {                                                                       \
    if ((L).timer) {                                                         \
      hipEvent_t ev_a_, ev_b_;                                               \
      (L).timer->events(cls, ev_a_, ev_b_);                                  \
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block),                   \
                            (std::uint32_t)(smem), (L).st, ev_a_, ev_b_, 0,  \
                            __VA_ARGS__);                                    \
    } else {                                                                 \
      kern<<<(grid), (block), (smem), (L).st>>>(__VA_ARGS__);                \
    }                                                                        \
  } while (0)
#define PBF_LAUNCH(L, cls, kern, grid, block, ...) \
  PBF_LAUNCH_SMEM(L, cls, kern, grid, block, 0, __VA_ARGS__)
// A tile kernel (kp: its persistent instance, k1: one work-group per tile)
// at work-group size bs in its class's launch form, with lds_ints ints of
// LDS per work-item (the replay row bases).
#define PBF_TILES(L, cls, bs, kp, k1, lds_ints, ...)                         \
  do {                                                                       \
    const int bs_ = (bs), nt_ = (L).sh.slots / bs_;                          \
    const size_t smem_ = (size_t)(lds_ints) * bs_ * sizeof(int);             \
    int g_ = (L).sh.grid[cls];                                               \
    if (g_ == kGridResident)                                                 \
      g_ = resident_grid(reinterpret_cast<const void *>(kp), bs_, smem_,     \
                         (L).sh.cus);                                        \
    if (g_ > 0 && g_ < nt_)                                                  \
      PBF_LAUNCH_SMEM(L, cls, kp, g_, bs_, smem_, __VA_ARGS__);              \
    else                                                                     \
      PBF_LAUNCH_SMEM(L, cls, k1, nt_, bs_, smem_, __VA_ARGS__);             \
  } while (0)
// A per-particle gather kernel at its class's work-group size.
#define PBF_GATHER(L, cls, kp, k1, lds_ints, ...) \
  PBF_TILES(L, cls, (L).sh.block[cls], kp, k1, lds_ints, __VA_ARGS__)

// cell_start from the cell counts, V int4 counts per work-item.
template <int V>
static void scan_stage(const DevBuffers &d, Launcher &L) {
  const StepShape &sh = L.sh;
  const int4 *count = reinterpret_cast<const int4 *>(d.count);
  int4 *cs = reinterpret_cast<int4 *>(d.cell_start);
  PBF_LAUNCH(L, kClsScan, k_scan_reduce<V>, sh.scan_blocks, 256, count,
             d.partial, d.dense);
  if (sh.scan_blocks <= kScanSumTiles) {
    PBF_LAUNCH(L, kClsScan, k_scan_down_sum<V>, sh.scan_blocks, 256, count,
               d.partial, cs);
  } else {
    PBF_LAUNCH(L, kClsScan, k_scan_top, 1, kTopThreads, d.partial,
               sh.scan_blocks);
    PBF_LAUNCH(L, kClsScan, k_scan_down<V>, sh.scan_blocks, 256, count,
               d.partial, cs);
  }
}

// The sort of a step.  with_predict: predict, bin and count here (probes,
// diagnostics); otherwise the previous step's last kernel (or
// enqueue_predict) has done it and the counts are in place.
template <int PH>
static void sort_stage(const DevBuffers &d, Launcher &L, int predict,
                       bool with_predict) {
  const StepShape &sh = L.sh;
  const int bs = 256;
  const int grid = sh.slots / bs;
  if (with_predict)
    PBF_LAUNCH(L, kClsPredict, k_predict_count<PH>, grid, bs, d.scenes,
               d.block_scene, d.pos, d.vel, d.pred_u, d.key, d.rank, d.count,
               predict);
  if (sh.scan_v == 4)
    scan_stage<4>(d, L);
  else if (sh.scan_v == 2)
    scan_stage<2>(d, L);
  else
    scan_stage<1>(d, L);
  PBF_TILES(L, kClsScatter, bs, (k_scatter<true>), (k_scatter<false>), 0,
            d.scenes, d.block_scene, d.key, d.rank, d.pid, d.cell_start,
            d.count, d.perm, d.ppid, d.dense, d.err);
  // x* as float4 (REF, and the probe and diagnostics of either physics)
  // and packed (PAPER)
  float4 *xa = PH == kRef || with_predict ? d.xa : nullptr;
  float *xp = PH == kPaper ? d.xp : nullptr;
  PBF_TILES(L, kClsReorder, bs, (k_reorder<true>), (k_reorder<false>), 0,
            d.scenes, d.block_scene, d.perm, d.ppid, d.key, d.cell_start,
            d.pred_u, d.pos, d.pid_s, xa, xp, d.xold, d.dense, d.dsc);
}

static NextSort next_sort(const DevBuffers &d) {
  return NextSort{d.pred_u, d.key, d.rank, d.count};
}

// A REF delta pass (RN: the shape's choice of IEEE sequences, ref_unit).
template <bool LAST>
static void delta_ref_pass(const DevBuffers &d, Launcher &L,
                           const float4 *cur, float4 *nxt) {
  if (L.sh.ref_rn)
    PBF_GATHER(L, kClsDelta, (k_delta_ref<true, LAST, true>),
               (k_delta_ref<false, LAST, true>), 0, d.scenes, d.block_scene,
               cur, d.cell_start, d.lam, d.rho, d.xold, nxt, d.v);
  else
    PBF_GATHER(L, kClsDelta, (k_delta_ref<true, LAST, false>),
               (k_delta_ref<false, LAST, false>), 0, d.scenes, d.block_scene,
               cur, d.cell_start, d.lam, d.rho, d.xold, nxt, d.v);
}

static void lambda_ref_pass(const DevBuffers &d, Launcher &L,
                            const float4 *cur) {
  if (L.sh.ref_rn)
    PBF_GATHER(L, kClsLambda, (k_lambda_ref<true, true>),
               (k_lambda_ref<false, true>), kRefQueue, d.scenes,
               d.block_scene, cur, d.cell_start, d.lam, d.rho);
  else
    PBF_GATHER(L, kClsLambda, (k_lambda_ref<true, false>),
               (k_lambda_ref<false, false>), kRefQueue, d.scenes,
               d.block_scene, cur, d.cell_start, d.lam, d.rho);
}

static void step_ref(const DevBuffers &d, Launcher &L) {
  const StepShape &sh = L.sh;
  sort_stage<kRef>(d, L, 1, false);
  float4 *cur = d.xa, *nxt = d.xb;
  for (int it = 0; it < sh.iterations; ++it) {
    lambda_ref_pass(d, L, cur);
    if (it == sh.iterations - 1)
      delta_ref_pass<true>(d, L, cur, nxt);
    else
      delta_ref_pass<false>(d, L, cur, nxt);
    std::swap(cur, nxt);
  }
  // XSPH's one reciprocal per pair: the range-tested sequence at every
  // size (c0: 121 against 125 us per launch with the general one)
  PBF_GATHER(L, kClsXsph, (k_xsph_final_ref<true, true>),
             (k_xsph_final_ref<false, true>), 0, d.scenes, d.block_scene, cur,
             d.cell_start, d.v, d.pid_s, d.pos, d.vel, d.pid, d.err,
             next_sort(d));
}

// A delta pass, at the s_corr exponent instance of the step's scenes.
template <bool LAST, bool WRITE_QA, bool WRITE_QV>
static void delta_pass(const DevBuffers &d, Launcher &L) {
  const int P = L.sh.slots;
  if (L.sh.scorr_n4)
    PBF_GATHER(L, kClsDelta,
               (k_delta_paper<true, LAST, WRITE_QA, true, WRITE_QV>),
               (k_delta_paper<false, LAST, WRITE_QA, true, WRITE_QV>), 9,
               d.scenes, d.block_scene, d.xp, d.xb, d.cell_start, d.rho,
               d.xold, d.nl, d.ncnt, d.rbase, P, d.xc, d.vr, d.qa, d.qv);
  else
    PBF_GATHER(L, kClsDelta,
               (k_delta_paper<true, LAST, WRITE_QA, false, WRITE_QV>),
               (k_delta_paper<false, LAST, WRITE_QA, false, WRITE_QV>), 9,
               d.scenes, d.block_scene, d.xp, d.xb, d.cell_start, d.rho,
               d.xold, d.nl, d.ncnt, d.rbase, P, d.xc, d.vr, d.qa, d.qv);
}

static void search_pass(const DevBuffers &d, Launcher &L) {
  PBF_GATHER(L, kClsSearch, (k_lambda_paper<true>), (k_lambda_paper<false>),
             0, d.scenes, d.block_scene, d.xp, d.cell_start, d.xb,
             d.rho, d.nl, d.ncnt, d.rbase, L.sh.slots);
}

// A lambda pass replaying the lists, over FP32 positions or (Q) over the
// 16-bit records.
template <bool Q>
static void lambda_list_pass(const DevBuffers &d, Launcher &L) {
  const int P = L.sh.slots;
  if constexpr (Q)
    PBF_GATHER(L, kClsLambda, (k_lambda_list_q<true>),
               (k_lambda_list_q<false>), 9, d.scenes, d.block_scene, d.xp,
               d.xc, d.qa, d.cell_start, d.nl, d.ncnt, d.rbase, P, d.xb,
               d.rho);
  else
    PBF_GATHER(L, kClsLambda, (k_lambda_list<true>), (k_lambda_list<false>),
               9, d.scenes, d.block_scene, d.xp, d.xc, d.cell_start, d.nl,
               d.ncnt, d.rbase, P, d.xb, d.rho);
}

// XSPH + omega and the vorticity pass (fused with the position update and
// the next step's predict), over FP32 records or (Q) the 16-bit ones.
template <bool Q>
static void vorticity_passes(const DevBuffers &d, Launcher &L) {
  const int P = L.sh.slots;
  if constexpr (Q)
    PBF_GATHER(L, kClsXsph, (k_xsph_omega_q<true>), (k_xsph_omega_q<false>),
               9, d.scenes, d.block_scene, d.xp, d.xc, d.vr, d.xold, d.qv,
               d.cell_start, d.nl, d.ncnt, d.rbase, P, d.vx, d.om, d.xb,
               d.qa);
  else
    PBF_GATHER(L, kClsXsph, (k_xsph_omega<true>), (k_xsph_omega<false>), 9,
               d.scenes, d.block_scene, d.xp, d.xc, d.vr, d.xold, d.qv,
               d.cell_start, d.nl, d.ncnt, d.rbase, P, d.vx, d.om, d.xb,
               d.qa);
  PBF_GATHER(L, kClsVort, (k_vort_final<true, Q>), (k_vort_final<false, Q>),
             9, d.scenes, d.block_scene, d.xp, d.xb, d.xc, d.qa, d.cell_start,
             d.vx, d.om, d.nl, d.ncnt, d.rbase, P, d.pid_s, d.pos, d.vel,
             d.pid, d.err, next_sort(d));
}

// XSPH fused with the position update and the next step's predict (no
// vorticity confinement).
template <bool Q>
static void xsph_final_pass(const DevBuffers &d, Launcher &L) {
  const int P = L.sh.slots;
  PBF_GATHER(L, kClsXsph, (k_xsph_final_paper<true, Q>),
             (k_xsph_final_paper<false, Q>), 9, d.scenes, d.block_scene, d.xp,
             d.xc, d.vr, d.xold, d.qv, d.cell_start, d.nl, d.ncnt, d.rbase, P,
             d.pid_s,
             d.pos, d.vel, d.pid, d.err, next_sort(d));
}

static void step_paper(const DevBuffers &d, Launcher &L) {
  const StepShape &sh = L.sh;
  sort_stage<kPaper>(d, L, 1, false);
  // mixed precision: the lambda passes of iterations 1 .. m - 1 (m =
  // fp16_iterations; iteration 0's is the FP32 search) gather 16-bit
  // positions, and so do XSPH and the vorticity pass; every delta pass
  // gathers FP32 records and updates FP32 positions
  const int m = sh.fp16_iterations;
  for (int it = 0; it < sh.iterations; ++it) {
    if (it == 0)
      search_pass(d, L);
    else if (it < m)
      lambda_list_pass<true>(d, L);
    else
      lambda_list_pass<false>(d, L);
    if (it == sh.iterations - 1) {
      if (m > 0)
        delta_pass<true, false, true>(d, L);
      else
        delta_pass<true, false, false>(d, L);
    } else if (it + 1 < m) {
      delta_pass<false, true, false>(d, L);
    } else {
      delta_pass<false, false, false>(d, L);
    }
  }
  if (sh.vorticity) {
    if (m > 0)
      vorticity_passes<true>(d, L);
    else
      vorticity_passes<false>(d, L);
  } else if (m > 0) {
    xsph_final_pass<true>(d, L);
  } else {
    xsph_final_pass<false>(d, L);
  }
}

void enqueue_step(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                  KernelTimer *timer) {
  Launcher L{sh, st, timer};
  if (sh.physics == kRef)
    step_ref(d, L);
  else
    step_paper(d, L);
}

// reps back-to-back launches of the representative kernel of class cls
// (search, lambda, delta), on the buffers a step left behind.  These
// kernels write only step temporaries (lists, (x, lambda), rho, x after a
// delta pass) that the next step rebuilds from the state before reading,
// and none writes what it reads, so the state is unchanged.  The values
// they compute from a previous pass's temporaries are not used.  Returns
// false for other classes.
bool repeatable_class(const StepShape &sh, int cls) {
  return cls == kClsLambda || cls == kClsDelta ||
         (cls == kClsSearch && sh.physics == kPaper);
}

bool enqueue_repeat(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                    int cls, int reps) {
  Launcher L{sh, st, nullptr};
  if (!repeatable_class(sh, cls)) return false;
  for (int r = 0; r < reps; ++r) {
    if (sh.physics == kRef) {
      if (cls == kClsLambda)
        lambda_ref_pass(d, L, d.xa);
      else
        delta_ref_pass<false>(d, L, d.xa, d.xb);
    } else if (cls == kClsSearch) {
      search_pass(d, L);
    } else if (cls == kClsLambda) {
      lambda_list_pass<false>(d, L);
    } else {
      delta_pass<false, false, false>(d, L);
    }
  }
  return true;
}

// Zero the cell counts (outside graphs only: a memset node is not captured).
static void clear_counts(const DevBuffers &d, const StepShape &sh,
                         hipStream_t st) {
  (void)hipMemsetAsync(d.count, 0,
                       (size_t)sh.scan_cells * sizeof(int), st);
}

void enqueue_predict(const DevBuffers &d, const StepShape &sh,
                     hipStream_t st) {
  clear_counts(d, sh, st);
  enqueue_predict_zeroed(d, sh, st);
}

void enqueue_predict_zeroed(const DevBuffers &d, const StepShape &sh,
                            hipStream_t st) {
  const int bs = 256, grid = sh.slots / bs;
  if (sh.physics == kRef)
    k_predict_count<kRef><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.pos,
                                               d.vel, d.pred_u, d.key, d.rank,
                                               d.count, 1);
  else
    k_predict_count<kPaper><<<grid, bs, 0, st>>>(
        d.scenes, d.block_scene, d.pos, d.vel, d.pred_u, d.key, d.rank,
        d.count, 1);
}

__global__ __launch_bounds__(256) void k_fill_int(int *__restrict__ a,
                                                  int value, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    a[i] = value;
}

__global__ void k_set_scene(Scene *__restrict__ out, Scene sc) { *out = sc; }

void enqueue_static_setup(const DevBuffers &d, const StepShape &sh,
                          const Scene &scene, hipStream_t st) {
  k_set_scene<<<1, 1, 0, st>>>(d.scenes, scene);
  auto fill = [&](int *a, int v, size_t n) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 1024);
    k_fill_int<<<std::max(grid, 1), 256, 0, st>>>(a, v, n);
  };
  fill(d.block_scene, 0, (size_t)sh.slots / kSlotAlign);
  fill(d.count, 0, (size_t)sh.scan_cells);
  fill(d.partial, 0, scan_partial_entries(sh));
  fill(d.err, 0, 1);
  fill(d.dense, 0, 2);
}

// Slot k < n takes particle k: x with .w = 0, v with .w = m, id k; the
// padding slots are zero with id -1.
__global__ __launch_bounds__(256) void k_load_state(
    const float4 *__restrict__ pos4, const float4 *__restrict__ vel4, int n,
    int slots, float4 *__restrict__ pos, float4 *__restrict__ vel,
    int *__restrict__ pid) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= slots) return;
  if (k < n) {
    const float4 x = pos4[k];
    pos[k] = make_float4(x.x, x.y, x.z, 0.0f);
    vel[k] = vel4[k];
    pid[k] = k;
  } else {
    pos[k] = vel[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pid[k] = -1;
  }
}

// The state of each slot back to its particle's place in original order.
__global__ __launch_bounds__(256) void k_store_state(
    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
    const int *__restrict__ pid, int n, float4 *__restrict__ pos4,
    float4 *__restrict__ vel4) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;  // the particles occupy the first n slots
  const int id = pid[k];
  if ((unsigned)id >= (unsigned)n) return;
  pos4[id] = pos[k];
  vel4[id] = vel[k];
}

// The float4 positions buffer a viewer draws, as the reference's
// updatePositions leaves it in the buffer it shares with GL
// (update_positions.cl:9-12): particle p's (x, |v|) at index p.  The scene's
// particles fill its first n slots (ids id_base .. id_base + n - 1).
__global__ __launch_bounds__(256) void k_store_positions(
    const float4 *__restrict__ pos, const int *__restrict__ pid, int n,
    int id_base, float4 *__restrict__ dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int id = pid[k] - id_base;
  if ((unsigned)id >= (unsigned)n) return;
  dst[id] = pos[k];
}

void enqueue_store_positions(const DevBuffers &d, const Scene &sc,
                             float4 *dst, hipStream_t st) {
  if (sc.n <= 0) return;
  k_store_positions<<<(sc.n + 255) / 256, 256, 0, st>>>(
      d.pos + sc.slot_off, d.pid + sc.slot_off, sc.n, sc.id_base, dst);
}

void enqueue_load_state(const DevBuffers &d, const StepShape &sh,
                        const float4 *pos4, const float4 *vel4, int n,
                        hipStream_t st) {
  k_load_state<<<sh.slots / 256, 256, 0, st>>>(pos4, vel4, n, sh.slots,
                                               d.pos, d.vel, d.pid);
}

void enqueue_store_state(const DevBuffers &d, const StepShape &sh,
                         float4 *pos4, float4 *vel4, int n, hipStream_t st) {
  k_store_state<<<(n + 255) / 256, 256, 0, st>>>(d.pos, d.vel, d.pid, n,
                                                 pos4, vel4);
}

void enqueue_probe(const DevBuffers &d, const StepShape &sh, hipStream_t st) {
  Launcher L{sh, st, nullptr};
  const int bs = 256, grid = sh.slots / bs;
  clear_counts(d, sh, st);
  if (sh.physics == kRef) {
    sort_stage<kRef>(d, L, 1, true);
    k_count_nbrs<kRef><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xa,
                                            d.cell_start, d.nbr);
    k_lambda_ref<false, false><<<grid, bs, kRefQueue * bs * sizeof(int), st>>>(
        d.scenes, d.block_scene, d.xa, d.cell_start, d.lam, d.rho);
  } else {
    sort_stage<kPaper>(d, L, 1, true);
    k_count_nbrs<kPaper><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xa,
                                              d.cell_start, d.nbr);
    k_lambda_paper<false><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xp,
                                               d.cell_start, d.xb, d.rho,
                                               d.nl, d.ncnt, d.rbase,
                                               sh.slots);
    k_unpack_lambda<<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xb,
                                         d.lam);
  }
}

void enqueue_diag(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                  double *out_dev) {
  Launcher L{sh, st, nullptr};
  const int bs = 256, grid = sh.slots / bs;
  clear_counts(d, sh, st);
  if (sh.physics == kRef) {
    sort_stage<kRef>(d, L, 0, true);
    k_diag<kRef><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xa,
                                      d.cell_start, d.vel, d.red);
  } else {
    sort_stage<kPaper>(d, L, 0, true);
    k_diag<kPaper><<<grid, bs, 0, st>>>(d.scenes, d.block_scene, d.xa,
                                        d.cell_start, d.vel, d.red);
  }
  k_diag_final<<<1, 256, 0, st>>>(d.red, grid, out_dev);
}

const char *kernel_class_name(int cls) {
  static const char *names[kNumClasses] = {
      "predict_count", "scan", "scatter",   "reorder", "lambda",
      "delta",         "xsph", "vorticity", "search"};
  return (cls >= 0 && cls < kNumClasses) ? names[cls] : "?";
}

}  // namespace pbf

namespace pbf {

// Plain streaming copy (measurement calibration only).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ a,
                                              float4 *__restrict__ b,
                                              size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    b[i] = a[i];
}

void enqueue_copy(const float4 *a, float4 *b, size_t n, hipStream_t st) {
  k_copy<<<2048, 256, 0, st>>>(a, b, n);
}

}  // namespace pbf
