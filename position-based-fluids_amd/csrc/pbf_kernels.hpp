// pbf_kernels.hpp - launch interface between the host runtime (pbf_api.cpp)
// and the HIP kernels (pbf_kernels.hip).
#pragma once

#include "pbf_scene.hpp"

namespace pbf {

constexpr int kRef = 0;
constexpr int kPaper = 1;

// Cell-count scan: 256 work-items x V int4 counts per work-group (a tile of
// kScanTile V cells, V = 1, 2 or 4: the cells per work-group the
// autotuner picks).  The count and cell_start arrays are allocated to a
// whole number of the largest tiles, at least M + 1 entries (the padding
// counts stay 0), so the scan needs no bounds tests at any V and writes
// cell_start[M] = total by itself.
constexpr int kScanTile = 1024;
constexpr int kScanMaxV = 4;
constexpr int kScanTileMax = kScanTile * kScanMaxV;
// Up to this many tiles, each tile of the down-sweep sums the tile totals
// before it itself (two launches); beyond, a one-work-group scan of the
// tile totals runs in between (three launches), since the summing work
// grows with the square of the tile count.
constexpr int kScanSumTiles = 4096;
// The one-work-group scan of the tile totals: kTopThreads work-items x
// kTopIpt totals per round; the totals array is allocated to whole rounds.
constexpr int kTopThreads = 1024;
constexpr int kTopIpt = 16;
constexpr int kTopChunk = kTopThreads * kTopIpt;

// Cells of more than kDenseCell slots are ranked by id over kDenseBuckets
// id buckets (the reorder's work-groups, once their tiles are done) instead
// of the reorder's count over the cell.
constexpr int kDenseCell = 256;
constexpr int kDenseBuckets = 1024;

// pbf_step_noalloc: bytes of the static device scratch (per GPU) that
// holds every temporary of the call (at 420 bytes per particle slot plus 8
// per cell: the 32,768-particle dam break in PAPER physics uses 14 MB).
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;

// Slots from which REF's pair terms use the range-tested IEEE sequences
// (the same bits as the general ones).
constexpr int kRefRnSlots = 1 << 18;

// StepShape::grid value of a persistent launch sized to the device's
// resident work-groups.
constexpr int kGridResident = -1;

// kernel classes reported by the profiler
enum KernelClass {
  kClsPredict = 0,
  kClsScan,
  kClsScatter,
  kClsReorder,
  kClsLambda,
  kClsDelta,
  kClsXsph,
  kClsVort,
  kClsSearch,  // PAPER: the lambda pass that searches and records N_i
  kNumClasses
};

// Shape of one step (fixed for a handle; captured into its graph).
struct StepShape {
  int slots;        // padded particle slots P (multiple of kSlotAlign)
  int cells;        // M, total cells over scenes
  int scan_cells;   // count / cell_start entries: M + 1 padded to a
                    // multiple of kScanTileMax
  int scan_v;       // int4 counts per work-item of the scan (1, 2 or 4)
  int scan_blocks;  // scan tiles: scan_cells / (kScanTile scan_v)
  int physics;      // kRef / kPaper (uniform over scenes)
  int iterations;
  int fp16_iterations;  // leading iterations gathering FP16 positions
  int vorticity;    // PAPER: vorticity confinement pass enabled
  int scorr_n4;     // every scene has the s_corr exponent n = 4
  int ref_rn;       // REF lambda and delta: range-tested shortened IEEE
                    // sequences (from kRefRnSlots slots: the step is
                    // throughput-bound) or the general ones (small steps,
                    // latency-bound)
  int block[kNumClasses];  // work-group size (64, 128 or 256) per class
  // launch form per class: 0 = one work-group per tile of block[c] slots;
  // g > 0 = persistent, g work-groups (a multiple of 8) walking the tiles;
  // kGridResident = persistent with the work-groups the device holds at
  // once (the kernel's occupancy times the compute units)
  int grid[kNumClasses];
  int cus;          // compute units of the device
};

// Scan tiles of scan_v int4 per work-item.
inline void set_scan_v(StepShape &sh, int v) {
  sh.scan_v = v;
  sh.scan_blocks = sh.scan_cells / (kScanTile * v);
}

// Entries of the scan's tile-total array: the tile count at scan_v 1,
// rounded up to whole rounds of the one-work-group scan.
inline size_t scan_partial_entries(const StepShape &sh) {
  return ((size_t)sh.scan_cells / kScanTile + kTopChunk) / kTopChunk *
         kTopChunk;
}

// Per-kernel-class event timing (pbf_profile_steps, the autotuner): hands
// out the start/stop event pair of the next launch of class cls.
struct KernelTimer {
  virtual void events(int cls, hipEvent_t &start, hipEvent_t &stop) = 0;
  virtual ~KernelTimer() = default;
};

// One step.  Its sort starts with the scan: the predicted positions, cell
// keys, arrival ranks and cell counts of the current state must be in place,
// left by the previous step's last kernel or by enqueue_predict, and the step
// leaves those of its result in place in turn.
void enqueue_step(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                  KernelTimer *timer);
// Zero the cell counts and predict, bin and count the current state (for
// the first step, and after anything that ran its own sort or replaced the
// state).  Not for graph capture.
void enqueue_predict(const DevBuffers &d, const StepShape &sh,
                     hipStream_t st);
bool repeatable_class(const StepShape &sh, int cls);
bool enqueue_repeat(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                    int cls, int reps);
// Probe and diagnostics sort the state themselves (zeroing the counts
// first) and leave the counts zeroed: enqueue_predict before the next step.
void enqueue_probe(const DevBuffers &d, const StepShape &sh, hipStream_t st);
void enqueue_diag(const DevBuffers &d, const StepShape &sh, hipStream_t st,
                  double *out_dev);
const char *kernel_class_name(int cls);
// Device address of the static scratch on the current device.
hipError_t static_scratch(void **ptr);
// pbf_step_noalloc's set-up without memset or copies: the scene table and
// block map filled by kernels, the counters the step needs at zero zeroed.
void enqueue_static_setup(const DevBuffers &d, const StepShape &sh,
                          const Scene &scene, hipStream_t st);
// Caller state (original order, n particles) into the slots, and back.
void enqueue_load_state(const DevBuffers &d, const StepShape &sh,
                        const float4 *pos4, const float4 *vel4, int n,
                        hipStream_t st);
void enqueue_store_state(const DevBuffers &d, const StepShape &sh,
                         float4 *pos4, float4 *vel4, int n, hipStream_t st);
// A viewer's positions buffer: (x, |v|) of each particle of scene sc at its
// index in input order, into dst (device memory, sc.n float4).
void enqueue_store_positions(const DevBuffers &d, const Scene &sc,
                             float4 *dst, hipStream_t st);
// enqueue_predict for counts already zeroed (no memset).
void enqueue_predict_zeroed(const DevBuffers &d, const StepShape &sh,
                            hipStream_t st);
void enqueue_copy(const float4 *a, float4 *b, size_t n, hipStream_t st);

}  // namespace pbf
