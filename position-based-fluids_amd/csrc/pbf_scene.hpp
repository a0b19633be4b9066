// pbf_scene.hpp - per-scene constants and device buffers shared by the HIP
// kernels (pbf_kernels.hip) and the host runtime (pbf_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pbf {

// Particle slots of every scene start at a multiple of this, so a
// work-group of <= kSlotAlign threads never straddles two scenes and the
// scene of a work-group is one scalar load.
constexpr int kSlotAlign = 256;

// PAPER binning: the x cells are kXSub times narrower than the y/z cells
// (Scene.nc[0] and cell_len[0] describe the fine x cells), so that the
// candidate range of a stencil row can be cut to the particle's reach.
constexpr int kXSub = 4;

// PAPER search reach: (1 + kReachMargin) h, a guard against rounding in the
// row cut (the search runs at the positions the bins were built from).
constexpr float kReachMargin = 1e-3f;

// Neighbour-list entries: the stencil row (0 .. 8, 4 bits) above kOffBits
// bits of offset from the row's first candidate slot (the first slot of the
// fine x cells the search visits in that row), kPerGroup entries packed in
// one 16-byte group from bit kGroupSkip up.  A row whose candidate range is
// longer than 2^kOffBits slots cannot be recorded (its particle walks its
// rows in every pass instead).
constexpr int kOffBits = 7;
constexpr int kEntryBits = 4 + kOffBits;
constexpr int kPerGroup = 128 / kEntryBits;
constexpr int kGroupSkip = 128 - kPerGroup * kEntryBits;

// Capacity of a particle's neighbour list (pairs within h at x*).
constexpr int kMaxGroups = 9;
constexpr int kMaxHits = kMaxGroups * kPerGroup;

// Mixed precision: neighbour positions as 16-bit fixed point in units of
// h / 2^kQBits from the scene's lower corner, kept modulo 2^16 (a period of
// 2^(16 - kQBits) h = 16 h): the difference of two such coordinates, taken
// modulo 2^16 as a signed 16-bit value, is the exact difference of the
// quantised positions for any two points less than 8 h apart (a listed pair
// is within h at x* and moves far less than 7 h within a step).
constexpr int kQBits = 12;

// Per-scene constants, resident in device memory and read by the kernels
// with scalar loads (the scene index is uniform per work-group).  REF
// scenes hold the 6-significant-digit constants of src/main.cpp:149-168.
struct Scene {
  float bmin[3], bmax[3];
  float cell_len[3];   // CELL_LENGTH_* (cell coordinate = (x-min)/len)
  float inv_cell[3];   // 1/len (PAPER cell coordinate = (x-min)*inv)
  int nc[3];           // cells per axis
  int cell_base;       // first global cell id of the scene
  int ncells;          // nc[0]*nc[1]*nc[2]
  int slot_off;        // first particle slot (multiple of kSlotAlign)
  int n;               // particles
  int shift;           // slot_off - (particles of the scenes before)
  int id_base;         // particles of the scenes before (= global id base)
  float h, h2, poly6, spiky;
  float dt, rho0, inv_rho0, eps;
  float c_xsph, g_y, scorr_k, scorr_cq;  // scorr_cq = poly6 / W(dq)
  int scorr_n;
  float vort_eps;
  float wave;          // x-min wall offset (Simulation::setWaveGenerator)
  float sphere[4];     // PAPER obstacle (cx, cy, cz, r), r <= 0: none
  float reach2;        // PAPER: ((1 + kReachMargin) h)^2
  float q_scale;       // mixed precision: 2^kQBits / h
  float q_unit;        // mixed precision: h / 2^kQBits
  int slots_all;       // padded slots of all scenes of the handle (the
                       // same in every scene; persistent launches)
};
static_assert(sizeof(Scene) % 16 == 0, "Scene must stay 16-byte aligned");

// A neighbour's record for XSPH as the gather returns it: x (from the packed
// positions) with 1/rho, and v, read from two arrays.
struct XVRec {
  float4 x;  // x, 1/rho (0: no density)
  float4 v;  // v, 0
};

// Device buffers of one handle.  Slot-indexed arrays have P entries (all
// scenes, padded); cell-indexed arrays M (+1) entries.  "Sorted" arrays are
// in (cell, particle id) order of the current step.
struct DevBuffers {
  // persistent state, sorted order of the previous step
  float4 *pos;    // x, y, z, |v|
  float4 *vel;    // v, m
  int *pid;       // global original particle id
  // step temporaries
  float4 *pred_u; // x*, m        (previous order)
  int *key;       // global cell id (previous order)
  int *rank;      // arrival rank in the cell (previous order)
  int *count;     // per cell, zero between steps
  int *cell_start;// M+1, exclusive scan of count (unpadded slots)
  int *partial;   // scan block sums
  int *perm;      // slot -> previous index, arrival order
  int *ppid;      // slot -> particle id, arrival order
  int *pid_s;     // sorted particle id
  int *dense;     // cells of more than kDenseCell slots, ranked apart:
                  // count, cursor, then (first slot, end) pairs
  int2 *dsc;      // their scratch (id, slot), one per slot
  float4 *xa;     // x* (sorted): bins, rows and neighbour lists of a step
  float *xp;      // PAPER: x* (sorted) as 3 floats per slot, the search's
                  // candidate positions (3 P + 16 floats)
  float4 *xb;     // (x, lambda) of a lambda pass (sorted)
  float *xc;      // PAPER: x after a delta pass but the last, packed (3 P
                  // + 16 floats), for the next lambda pass
  float4 *xold;   // x at step start, m (sorted)
  float *lam;     // lambda (sorted)
  float *rho;     // density of the last lambda pass (sorted)
  float4 *v;      // REF: velocity after the constraint solve (sorted), .w = m
  float4 *vr;     // PAPER: (v, 1/rho) after the last delta pass (its x in
                  // xc), for XSPH
  float4 *vx;     // PAPER: v + XSPH (sorted), .w = m
  float4 *om;     // PAPER: vorticity, .w = |omega| (sorted)
  // mixed precision records (sorted), 16-bit fixed-point x (x | y << 16,
  // z | payload << 16) and FP16 payloads:
  uint2 *qa;      //   x after a delta pass (payload 0), for a lambda pass;
                  //   x and |omega| after XSPH, for the vorticity pass
  uint4 *qv;      //   x, rho0 / rho, v after the last delta pass, for XSPH
  uint4 *nl;      // PAPER neighbour lists: kMaxGroups groups of kPerGroup
                  // entries per slot, group g of slot i at [g * P + i]
  int *ncnt;      // list length per slot (> kMaxHits: list overflowed)
  int *rbase;     // PAPER: first candidate slot of each stencil row (the
                  // base of its list entries), [row * P + i]
  int *nbr;       // probe: neighbour counts (sorted)
  double *red;    // diagnostics partials
  int *err;       // device error word
  int *block_scene;  // per kSlotAlign block of slots -> scene
  Scene *scenes;
};

}  // namespace pbf
