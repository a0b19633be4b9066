// pbf_api.cpp - host runtime behind the C ABI of include/pbf.h.
//
// Replaces the reference's Simulation + OpenCL plumbing
// (src/Simulation.cpp, src/ocl/clsetup.cpp): owns the device state of one
// or many scenes, sequences a step as one captured hipGraph on the handle's
// stream, and converts every failure into a status code plus a
// thread-local message (no exception crosses the ABI).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_kernels.hpp"

namespace {

thread_local std::string g_error;

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define PBF_HIP(call)                                                     \
  do {                                                                    \
    const hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess)                                                 \
      throw ApiError(PBF_ERR_HIP, std::string(#call) + ": " +             \
                                      hipGetErrorString(e_));             \
  } while (0)

[[noreturn]] void invalid(const std::string &m) {
  throw ApiError(PBF_ERR_INVALID, m);
}

template <typename F>
int guard(F &&f) {
  try {
    f();
    g_error.clear();
    return PBF_OK;
  } catch (const ApiError &e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_error = "host allocation failed";
    return PBF_ERR_HIP;
  } catch (const std::exception &e) {
    g_error = e.what();
    return PBF_ERR_HIP;
  }
}

// 6-significant-digit rendering of a kernel constant: src/main.cpp:149-168
// streams each value with std::showpoint at the default precision into a
// -D flag, i.e. printf "%#g".
float fmt6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%#g", v);
  return std::strtof(buf, nullptr);
}

bool is_int_valued(float v) { return std::floor(v) == v; }

void validate(const pbf_params &p) {
  if (!(p.x_n >= 1 && p.y_n >= 1 && p.z_n >= 1) || !is_int_valued(p.x_n) ||
      !is_int_valued(p.y_n) || !is_int_valued(p.z_n))
    invalid("cell counts x_n, y_n, z_n must be integers >= 1");
  if (!(p.x_max > p.x_min && p.y_max > p.y_min && p.z_max > p.z_min))
    invalid("box must have max > min on every axis");
  // PAPER bins x kXSub times finer; cell ids are 32-bit
  const double xsub = p.physics == PBF_PHYSICS_PAPER ? pbf::kXSub : 1;
  if ((double)p.x_n * xsub * p.y_n * p.z_n > (double)(1 << 30))
    invalid("too many cells (limit 2^30, PAPER x cells counted 4 times)");
  if (!(p.time_step_length > 0.0f)) invalid("timestep_length must be > 0");
  if (!(p.rest_density > 0.0f)) invalid("restdensity must be > 0");
  if (p.physics != PBF_PHYSICS_REF && p.physics != PBF_PHYSICS_PAPER)
    invalid("physics must be PBF_PHYSICS_REF or PBF_PHYSICS_PAPER");
  if (p.iterations < 1 || p.iterations > 64)
    invalid("iterations must be in [1, 64]");
  if (p.precision != PBF_PRECISION_FP32 && p.precision != PBF_PRECISION_MIXED)
    invalid("precision must be PBF_PRECISION_FP32 or PBF_PRECISION_MIXED");
  if (p.precision == PBF_PRECISION_MIXED &&
      (p.fp32_iterations < 1 || p.fp32_iterations > p.iterations))
    invalid("mixed precision needs 1 <= fp32_iterations <= iterations");
  if (p.precision == PBF_PRECISION_MIXED && p.physics != PBF_PHYSICS_PAPER)
    invalid("mixed precision is available for PAPER physics only");
  if (p.physics == PBF_PHYSICS_PAPER) {
    // the 3 x 3 rows of the search need y and z cells at least h = x cell
    const double hx = ((double)p.x_max - p.x_min) / p.x_n;
    const double cy = ((double)p.y_max - p.y_min) / p.y_n;
    const double cz = ((double)p.z_max - p.z_min) / p.z_n;
    if (cy < hx * (1.0 - 1e-6) || cz < hx * (1.0 - 1e-6))
      invalid("PAPER physics needs y and z cells at least as long as the "
              "x cells (h)");
  }
  if (!(p.scorr_n >= 0 && p.scorr_n <= 16) || !is_int_valued(p.scorr_n))
    invalid("scorr_n must be an integer in [0, 16]");
  if (p.block_size != 0 && p.block_size != 64 && p.block_size != 128 &&
      p.block_size != 256)
    invalid("block_size must be 0, 64, 128 or 256");
  if (p.launch_grid != -1 &&
      (p.launch_grid < 0 || p.launch_grid % 8 != 0 ||
       p.launch_grid > (1 << 20)))
    invalid("launch_grid must be -1, 0 or a multiple of 8 up to 2^20");
  if (p.scan_cells != 0 && p.scan_cells != 1024 && p.scan_cells != 2048 &&
      p.scan_cells != 4096)
    invalid("scan_cells must be 0, 1024, 2048 or 4096");
}

pbf::Scene make_scene(const pbf_params &p) {
  pbf::Scene s{};
  const bool ref = p.physics == PBF_PHYSICS_REF;
  const float bmin[3] = {p.x_min, p.y_min, p.z_min};
  const float bmax[3] = {p.x_max, p.y_max, p.z_max};
  const float n[3] = {p.x_n, p.y_n, p.z_n};
  const float hf = (p.x_max - p.x_min) / p.x_n;  // main.cpp:164
  for (int a = 0; a < 3; ++a) {
    const float cl = (bmax[a] - bmin[a]) / n[a];  // main.cpp:159-161
    if (ref) {
      s.bmin[a] = fmt6(bmin[a]);
      s.bmax[a] = fmt6(bmax[a]);
      s.cell_len[a] = fmt6(cl);
    } else {
      s.bmin[a] = bmin[a];
      s.bmax[a] = bmax[a];
      s.cell_len[a] = (float)(((double)bmax[a] - (double)bmin[a]) / n[a]);
    }
    s.nc[a] = (int)n[a];
    if (!ref && a == 0) {  // fine x bins of the PAPER search
      s.nc[0] *= pbf::kXSub;
      s.cell_len[0] = (float)(((double)bmax[0] - (double)bmin[0]) /
                              ((double)n[0] * pbf::kXSub));
    }
    s.inv_cell[a] = 1.0f / s.cell_len[a];
  }
  s.ncells = s.nc[0] * s.nc[1] * s.nc[2];
  const double pi = 3.14159265358979323846;
  double h;
  if (ref) {
    h = (double)hf;
    s.h = fmt6(hf);
    s.h2 = fmt6(std::pow((double)hf, 2));
    s.poly6 = fmt6(315.0 / (64.0 * pi * std::pow((double)hf, 9)));
    s.spiky = fmt6(45.0 / (pi * std::pow((double)hf, 6)));
    s.dt = fmt6(p.time_step_length);
    s.rho0 = fmt6(p.rest_density);
  } else {
    h = ((double)p.x_max - (double)p.x_min) / (double)p.x_n;
    s.h = (float)h;
    s.h2 = (float)(h * h);
    s.poly6 = (float)(315.0 / (64.0 * pi * std::pow(h, 9)));
    s.spiky = (float)(45.0 / (pi * std::pow(h, 6)));
    s.dt = p.time_step_length;
    s.rho0 = p.rest_density;
    const double reach = h * (1.0 + (double)pbf::kReachMargin);
    s.reach2 = (float)(reach * reach);
    s.q_scale = (float)((double)(1 << pbf::kQBits) / h);
    s.q_unit = (float)(h / (double)(1 << pbf::kQBits));
  }
  s.inv_rho0 = 1.0f / s.rho0;
  s.eps = p.relaxation;
  s.c_xsph = p.xsph_c;
  s.g_y = p.gravity_y;
  s.scorr_k = p.scorr_k;
  s.scorr_n = (int)p.scorr_n;
  {
    const double hh = (double)s.h, dq = (double)p.scorr_dq * hh;
    const double t = hh * hh - dq * dq;
    s.scorr_cq = (t > 0.0) ? (float)(1.0 / (t * t * t)) : 0.0f;
  }
  s.vort_eps = p.vorticity_eps;
  s.wave = 0.0f;
  for (int a = 0; a < 4; ++a) s.sphere[a] = p.sphere[a];
  return s;
}

struct EventTimer final : pbf::KernelTimer {
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t take() {
    if (used == pool.size()) {
      hipEvent_t e;
      PBF_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void events(int cls, hipEvent_t &start, hipEvent_t &stop) override {
    Rec r{cls, take(), nullptr};
    r.b = take();
    recs.push_back(r);
    start = r.a;
    stop = r.b;
  }
  ~EventTimer() override {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

namespace pbf {
// Error text of the scenario readers (io_capi.cpp).
void set_error_text(const char *text) { g_error = text ? text : ""; }
}  // namespace pbf

struct pbf_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<pbf_params> params;
  std::vector<pbf::Scene> scenes;
  pbf::Scene *scenes_pinned = nullptr;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  bool dirty = true;
  std::vector<int> block_scene;
  pbf::StepShape shape{};
  int total = 0;  // particles over scenes
  pbf::DevBuffers d{};
  std::vector<void *> allocs;
  hipGraphExec_t graph = nullptr;        // one step
  hipGraphExec_t graph_multi = nullptr;  // kGraphSteps steps
  double *diag_out = nullptr;
  float wave = 0.0f;
  bool tune_pending = false;  // block size still to be chosen
  // the predicted positions, keys, ranks and counts of the current state
  // are in place (a step needs them; probes and the autotuner discard them)
  bool predicted = false;

  template <typename T>
  T *alloc(size_t count) {
    void *p = nullptr;
    PBF_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    allocs.push_back(p);
    return static_cast<T *>(p);
  }

  ~pbf_sim() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_multi) (void)hipGraphExecDestroy(graph_multi);
    for (void *p : allocs) (void)hipFree(p);
    if (scenes_pinned) (void)hipHostFree(scenes_pinned);
    if (upload_done) (void)hipEventDestroy(upload_done);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // Copy the scene table to the device if it changed (wave generator).
  void upload_scenes() {
    if (!dirty) return;
    if (upload_pending) PBF_HIP(hipEventSynchronize(upload_done));
    std::memcpy(scenes_pinned, scenes.data(),
                scenes.size() * sizeof(pbf::Scene));
    PBF_HIP(hipMemcpyAsync(d.scenes, scenes_pinned,
                           scenes.size() * sizeof(pbf::Scene),
                           hipMemcpyHostToDevice, stream));
    PBF_HIP(hipEventRecord(upload_done, stream));
    upload_pending = true;
    dirty = false;
  }

  void check_err() {
    int e = 0;
    PBF_HIP(hipMemcpyAsync(&e, d.err, sizeof(int), hipMemcpyDeviceToHost,
                           stream));
    PBF_HIP(hipStreamSynchronize(stream));
    if (e & 2)
      throw ApiError(PBF_ERR_STATE,
                     "internal error: cell counts disagree with cell keys");
    if (e)
      throw ApiError(PBF_ERR_NUMERIC,
                     "non-finite particle position or velocity");
  }

  void ensure_predicted() {
    if (predicted) return;
    pbf::enqueue_predict(d, shape, stream);
    predicted = true;
  }
};

namespace {

// Shape of a step over scenes of `scenes` (laid out, slot_off and cell_base
// set) with the uniform settings of p0; work-group sizes untuned.
pbf::StepShape make_shape(const pbf_params &p0, const pbf::Scene *scenes,
                          size_t nscenes, const pbf_params *params,
                          long long slots, long long cells) {
  pbf::StepShape sh{};
  sh.slots = (int)slots;
  sh.cells = (int)cells;
  sh.scan_cells = (int)((cells + pbf::kScanTileMax) / pbf::kScanTileMax *
                         pbf::kScanTileMax);
  pbf::set_scan_v(sh, p0.scan_cells ? p0.scan_cells / pbf::kScanTile : 1);
  sh.physics = p0.physics == PBF_PHYSICS_REF ? pbf::kRef : pbf::kPaper;
  sh.iterations = p0.iterations;
  // the lambda passes of iterations 1 .. I - F - 1, XSPH and the vorticity
  // pass gather 16-bit records; iteration 0's search, every delta pass and
  // the last F iterations are FP32
  sh.fp16_iterations = p0.precision == PBF_PRECISION_MIXED
                           ? std::max(p0.iterations - p0.fp32_iterations, 0)
                           : 0;
  bool vort = false;
  for (size_t s = 0; s < nscenes; ++s)
    vort = vort || params[s].vorticity_eps != 0.0f;
  sh.vorticity = (sh.physics == pbf::kPaper && vort) ? 1 : 0;
  sh.scorr_n4 = 1;
  for (size_t s = 0; s < nscenes; ++s) sh.scorr_n4 &= scenes[s].scorr_n == 4;
  sh.ref_rn = slots >= pbf::kRefRnSlots ? 1 : 0;
  for (int c = 0; c < pbf::kNumClasses; ++c) {
    sh.block[c] = p0.block_size ? p0.block_size : 128;
    sh.grid[c] = p0.launch_grid;
  }
  return sh;
}

// Every device buffer of a step of shape sh over nscenes scenes, placed by
// get(bytes) (an allocation, or a bump pointer into preallocated memory).
template <typename Get>
void layout_buffers(pbf::DevBuffers &d, const pbf::StepShape &sh,
                    size_t nscenes, Get &&get) {
  auto take = [&](auto *&ptr, size_t count) {
    using T = std::remove_reference_t<decltype(*ptr)>;
    ptr = static_cast<T *>(get(std::max<size_t>(count, 1) * sizeof(T)));
  };
  const size_t P = (size_t)sh.slots;
  take(d.pos, P);
  take(d.vel, P);
  take(d.pid, P);
  take(d.pred_u, P);
  take(d.key, P);
  take(d.rank, P);
  const size_t Mpad = (size_t)sh.scan_cells;  // >= M + 1
  take(d.count, Mpad);
  take(d.cell_start, Mpad);
  take(d.partial, pbf::scan_partial_entries(sh));
  take(d.perm, P);
  take(d.ppid, P);
  take(d.pid_s, P);
  take(d.dense, 2 + 2 * (P / pbf::kDenseCell + 1));
  take(d.dsc, P);
  take(d.xa, P);
  take(d.xb, P);
  if (sh.physics == pbf::kPaper) {
    take(d.xp, 3 * P + 16);
    take(d.xc, 3 * P + 16);
    take(d.vr, P);
  }
  take(d.xold, P);
  take(d.lam, P);
  take(d.rho, P);
  if (sh.physics == pbf::kRef) take(d.v, P);
  if (sh.vorticity) {
    take(d.vx, P);
    take(d.om, P);
  }
  if (sh.fp16_iterations > 0) {
    take(d.qa, P);
    take(d.qv, P);
  }
  if (sh.physics == pbf::kPaper) {
    take(d.nl, P * (size_t)pbf::kMaxGroups);
    take(d.ncnt, P);
    take(d.rbase, P * 9);
  }
  take(d.nbr, P);
  take(d.red, (P / 64 + 1) * 8);
  take(d.err, 1);
  take(d.block_scene, P / pbf::kSlotAlign);
  take(d.scenes, nscenes);
}

void build(pbf_sim &S, uint32_t nscenes, const pbf_params *params,
           const pbf_particle *const *parts, const uint32_t *counts) {
  if (nscenes == 0) invalid("nscenes must be >= 1");
  for (uint32_t s = 0; s < nscenes; ++s) {
    validate(params[s]);
    if (counts[s] > 0 && !parts[s]) invalid("particles is NULL");
    if (params[s].physics != params[0].physics ||
        params[s].iterations != params[0].iterations ||
        params[s].precision != params[0].precision ||
        params[s].fp32_iterations != params[0].fp32_iterations)
      invalid("all scenes of a batch need the same physics, iterations and "
              "precision");
  }
  PBF_HIP(hipGetDevice(&S.device));
  PBF_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  PBF_HIP(hipEventCreateWithFlags(&S.upload_done, hipEventDisableTiming));
  S.params.assign(params, params + nscenes);
  long long slot = 0, cells = 0, ids = 0;
  for (uint32_t s = 0; s < nscenes; ++s) {
    pbf::Scene sc = make_scene(params[s]);
    sc.n = (int)counts[s];
    sc.slot_off = (int)slot;
    sc.id_base = (int)ids;
    sc.shift = (int)(slot - ids);
    sc.cell_base = (int)cells;
    const long long padded =
        ((long long)counts[s] + pbf::kSlotAlign - 1) / pbf::kSlotAlign *
        pbf::kSlotAlign;
    for (long long b = 0; b < padded / pbf::kSlotAlign; ++b)
      S.block_scene.push_back((int)s);
    slot += padded;
    ids += counts[s];
    cells += sc.ncells;
    if (slot > (1LL << 30) || cells > (1LL << 30))
      invalid("problem too large (slots or cells beyond 2^30)");
    S.scenes.push_back(sc);
  }
  S.total = (int)ids;
  for (pbf::Scene &sc : S.scenes) sc.slots_all = (int)slot;
  // PAPER gathers address arrays of up to 16-byte records with 32-bit byte
  // offsets
  if (params[0].physics == PBF_PHYSICS_PAPER && slot > INT32_MAX / 16)
    invalid("problem too large for PAPER physics (more than 2^27 slots)");
  const pbf_params &p0 = params[0];
  pbf::StepShape &sh = S.shape;
  sh = make_shape(p0, S.scenes.data(), S.scenes.size(), params, slot, cells);
  PBF_HIP(hipDeviceGetAttribute(&sh.cus, hipDeviceAttributeMultiprocessorCount,
                                S.device));
  S.tune_pending =
      p0.block_size == 0 && p0.launch_grid == 0 && p0.scan_cells == 0;

  const size_t P = (size_t)sh.slots;
  pbf::DevBuffers &d = S.d;
  layout_buffers(d, sh, nscenes, [&](size_t bytes) {
    return static_cast<void *>(S.alloc<char>(bytes));
  });
  const size_t Mpad = (size_t)sh.scan_cells;
  const size_t Tpad = pbf::scan_partial_entries(sh);
  S.diag_out = S.alloc<double>(8);
  PBF_HIP(hipHostMalloc((void **)&S.scenes_pinned,
                        nscenes * sizeof(pbf::Scene), hipHostMallocDefault));

  // Pack the particles as Simulation::init does (Simulation.cpp:120-134):
  // pos = (x, 0), vel = (v, m); slots start in input order.
  std::vector<float4> hpos(P, make_float4(0, 0, 0, 0)), hvel(P, hpos[0]);
  std::vector<int> hpid(P, -1);
  for (uint32_t s = 0; s < nscenes; ++s) {
    const pbf::Scene &sc = S.scenes[s];
    for (uint32_t k = 0; k < counts[s]; ++k) {
      const pbf_particle &q = parts[s][k];
      hpos[sc.slot_off + k] = make_float4(q.x[0], q.x[1], q.x[2], 0.0f);
      hvel[sc.slot_off + k] = make_float4(q.v[0], q.v[1], q.v[2], q.m);
      hpid[sc.slot_off + k] = sc.id_base + (int)k;
    }
  }
  PBF_HIP(hipMemcpy(d.pos, hpos.data(), P * sizeof(float4),
                    hipMemcpyHostToDevice));
  PBF_HIP(hipMemcpy(d.vel, hvel.data(), P * sizeof(float4),
                    hipMemcpyHostToDevice));
  PBF_HIP(hipMemcpy(d.pid, hpid.data(), P * sizeof(int),
                    hipMemcpyHostToDevice));
  PBF_HIP(hipMemset(d.count, 0, Mpad * sizeof(int)));
  PBF_HIP(hipMemset(d.partial, 0, Tpad * sizeof(int)));
  PBF_HIP(hipMemset(d.err, 0, sizeof(int)));
  PBF_HIP(hipMemset(d.dense, 0, 2 * sizeof(int)));
  PBF_HIP(hipMemcpy(d.block_scene, S.block_scene.data(),
                    S.block_scene.size() * sizeof(int),
                    hipMemcpyHostToDevice));
  S.upload_scenes();
  PBF_HIP(hipStreamSynchronize(S.stream));
}

// ---- autotuner ---------------------------------------------------------
// Chooses the launch of each kernel class: work-group size 64, 128 or 256,
// each with one work-group per tile or persistent (as many work-groups as
// the device holds at once, walking all tiles), by timing each class over
// two steps in every combination, from the handle's own state advanced 10
// steps (the state is saved before and restored after); the classes are
// independent launches, so the fastest combination of each is taken (the
// default unless another beats it by more than 2 %).  Then the cells per
// work-group of the cell-count scan: 1024, 2048 or 4096.  Results are kept
// per problem shape for the process and, when PBF_TUNE_CACHE names a file,
// in that file ("key b g b g ... v" lines: size and grid per class, then
// the scan's int4 counts per work-item).  Handles below kTuneMinSlots slots
// keep the default: there the launch overhead, not the kernels, dominates.
constexpr int kTuneMinSlots = 1 << 18;
struct Launch {
  std::array<int, pbf::kNumClasses> block, grid;
  int scan_v = 1;
};

std::mutex g_tune_mu;
std::map<std::string, Launch> g_tune;

std::string tune_key(const pbf_sim &S) {
  hipDeviceProp_t prop;
  std::string arch = "gpu";
  if (hipGetDeviceProperties(&prop, S.device) == hipSuccess)
    arch = prop.gcnArchName;
  std::ostringstream k;
  k << arch << ":" << S.shape.slots << ":" << S.shape.cells << ":"
    << S.shape.physics << ":" << S.shape.iterations << ":"
    << S.shape.fp16_iterations << ":" << S.shape.vorticity << ":"
    << S.scenes.size() << ":c" << pbf::kNumClasses << ":t3";
  return k.str();
}

bool valid_block(int b) { return b == 64 || b == 128 || b == 256; }
bool valid_grid(int g) {
  return g == pbf::kGridResident || (g >= 0 && g % 8 == 0 && g <= (1 << 20));
}

bool tune_lookup(const std::string &key, Launch &out) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    out = it->second;
    return true;
  }
  if (const char *path = std::getenv("PBF_TUNE_CACHE")) {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ls(line);
      std::string k;
      Launch l;
      if (!(ls >> k) || k != key) continue;
      bool ok = true;
      for (int c = 0; c < pbf::kNumClasses; ++c)
        ok = ok && (ls >> l.block[c] >> l.grid[c]) && valid_block(l.block[c]) &&
             valid_grid(l.grid[c]);
      ok = ok && (ls >> l.scan_v) &&
           (l.scan_v == 1 || l.scan_v == 2 || l.scan_v == 4);
      if (ok) {
        g_tune[key] = l;
        out = l;
        return true;
      }
    }
  }
  return false;
}

void tune_store(const std::string &key, const Launch &l) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[key] = l;
  if (const char *path = std::getenv("PBF_TUNE_CACHE")) {
    std::ofstream f(path, std::ios::app);
    f << key;
    for (int c = 0; c < pbf::kNumClasses; ++c)
      f << " " << l.block[c] << " " << l.grid[c];
    f << " " << l.scan_v << "\n";
  }
}

void apply_launch(pbf::StepShape &sh, const Launch &l) {
  std::copy(l.block.begin(), l.block.end(), sh.block);
  std::copy(l.grid.begin(), l.grid.end(), sh.grid);
  pbf::set_scan_v(sh, l.scan_v);
}

void autotune(pbf_sim &S) {
  if (!S.tune_pending) return;
  S.tune_pending = false;
  if (S.shape.slots < kTuneMinSlots) return;
  const std::string key = tune_key(S);
  Launch cached;
  if (tune_lookup(key, cached)) {
    apply_launch(S.shape, cached);
    return;
  }
  const size_t P = (size_t)S.shape.slots;
  // the handle's state (restored at the end) and the trials' start
  float4 *opos = nullptr, *ovel = nullptr, *spos = nullptr, *svel = nullptr;
  int *opid = nullptr, *spid = nullptr;
  auto release = [&] {
    for (void *p : {(void *)opos, (void *)ovel, (void *)opid, (void *)spos,
                    (void *)svel, (void *)spid})
      if (p) (void)hipFree(p);
  };
  auto copy = [&](void *dst, const void *src, size_t bytes) {
    PBF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S.stream));
  };
  try {
    PBF_HIP(hipMalloc(&opos, P * sizeof(float4)));
    PBF_HIP(hipMalloc(&ovel, P * sizeof(float4)));
    PBF_HIP(hipMalloc(&opid, P * sizeof(int)));
    PBF_HIP(hipMalloc(&spos, P * sizeof(float4)));
    PBF_HIP(hipMalloc(&svel, P * sizeof(float4)));
    PBF_HIP(hipMalloc(&spid, P * sizeof(int)));
    copy(opos, S.d.pos, P * sizeof(float4));
    copy(ovel, S.d.vel, P * sizeof(float4));
    copy(opid, S.d.pid, P * sizeof(int));
    // the trials start kTuneLead steps into the run (at the default
    // launches): a scene's first steps, often a lattice at rest, are not
    // the ones its run spends its time in
    constexpr int kTuneLead = 10;
    pbf::enqueue_predict(S.d, S.shape, S.stream);
    for (int t = 0; t < kTuneLead; ++t)
      pbf::enqueue_step(S.d, S.shape, S.stream, nullptr);
    copy(spos, S.d.pos, P * sizeof(float4));
    copy(svel, S.d.vel, P * sizeof(float4));
    copy(spid, S.d.pid, P * sizeof(int));
    Launch best;
    best.block.fill(S.shape.block[0]);
    best.grid.fill(0);
    std::array<double, pbf::kNumClasses> best_ms;
    best_ms.fill(1e30);
    // two timed steps of the current shape from the saved state: ms per
    // kernel class
    auto trial = [&] {
      copy(S.d.pos, spos, P * sizeof(float4));
      copy(S.d.vel, svel, P * sizeof(float4));
      copy(S.d.pid, spid, P * sizeof(int));
      pbf::enqueue_predict(S.d, S.shape, S.stream);
      pbf::enqueue_step(S.d, S.shape, S.stream, nullptr);  // warm-up
      EventTimer timer;
      for (int t = 0; t < 2; ++t)
        pbf::enqueue_step(S.d, S.shape, S.stream, &timer);
      PBF_HIP(hipStreamSynchronize(S.stream));
      std::array<double, pbf::kNumClasses> ms{};
      for (const auto &r : timer.recs) {
        float t = 0.0f;
        PBF_HIP(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.cls] += t;
      }
      return ms;
    };
    // A choice other than the default (128 work-items, one work-group per
    // tile, 1024-cell scan tiles) is kept only when it is faster by more
    // than kTuneMargin: the trials time the handle's first steps, and
    // near-ties there resolve at random (and need not hold later in a run).
    constexpr double kTuneMargin = 0.02;
    std::array<double, pbf::kNumClasses> def_ms;
    def_ms.fill(1e30);
    pbf::set_scan_v(S.shape, 1);
    for (int b : {64, 128, 256}) {
      // one work-group per tile, then persistent with the work-groups the
      // device holds at once
      for (int g : {0, pbf::kGridResident}) {
        for (int c = 0; c < pbf::kNumClasses; ++c) {
          S.shape.block[c] = b;
          S.shape.grid[c] = g;
        }
        const auto ms = trial();
        for (int c = 0; c < pbf::kNumClasses; ++c) {
          if (b == 128 && g == 0) def_ms[c] = ms[c];
          if (ms[c] < best_ms[c]) {
            best_ms[c] = ms[c];
            best.block[c] = b;
            best.grid[c] = g;
          }
        }
      }
    }
    for (int c = 0; c < pbf::kNumClasses; ++c)
      if (def_ms[c] <= best_ms[c] * (1.0 + kTuneMargin)) {
        best.block[c] = 128;
        best.grid[c] = 0;
      }
    // the scan's tile at the chosen launches (1024 cells per work-group
    // timed with the rest)
    apply_launch(S.shape, best);
    const double scan_def = def_ms[pbf::kClsScan];
    double scan_best = scan_def;
    for (int v : {2, 4}) {
      pbf::set_scan_v(S.shape, v);
      const auto ms = trial();
      if (ms[pbf::kClsScan] < scan_best) {
        scan_best = ms[pbf::kClsScan];
        best.scan_v = v;
      }
    }
    if (scan_def <= scan_best * (1.0 + kTuneMargin)) best.scan_v = 1;
    copy(S.d.pos, opos, P * sizeof(float4));
    copy(S.d.vel, ovel, P * sizeof(float4));
    copy(S.d.pid, opid, P * sizeof(int));
    S.predicted = false;  // the trial steps left their own
    S.ensure_predicted();
    PBF_HIP(hipStreamSynchronize(S.stream));
    apply_launch(S.shape, best);
    tune_store(key, best);
  } catch (...) {
    S.predicted = false;
    release();
    throw;
  }
  release();
}

// Steps per launch of the multi-step graph: a graph launch leaves a gap of
// about 8 us on the GPU before its first kernel (c1 kernel trace), 1 % of a
// 1M-particle step, paid once per kGraphSteps steps instead of every step.
constexpr uint32_t kGraphSteps = 8;

// A graph of `steps` steps of the handle, captured and instantiated.
hipGraphExec_t capture_steps(pbf_sim &S, uint32_t steps) {
  hipGraph_t g = nullptr;
  PBF_HIP(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
  for (uint32_t t = 0; t < steps; ++t)
    pbf::enqueue_step(S.d, S.shape, S.stream, nullptr);
  PBF_HIP(hipStreamEndCapture(S.stream, &g));
  hipGraphExec_t ge = nullptr;
  const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    throw ApiError(PBF_ERR_HIP, std::string("hipGraphInstantiate: ") +
                                    hipGetErrorString(e));
  return ge;
}

void ensure_graph(pbf_sim &S) {
  autotune(S);
  if (!S.graph) S.graph = capture_steps(S, 1);
  if (!S.graph_multi) S.graph_multi = capture_steps(S, kGraphSteps);
}

// nsteps steps as graph launches: kGraphSteps at a time, then one at a time.
void launch_steps(pbf_sim &S, uint32_t nsteps) {
  for (; nsteps >= kGraphSteps; nsteps -= kGraphSteps)
    PBF_HIP(hipGraphLaunch(S.graph_multi, S.stream));
  for (; nsteps > 0; --nsteps) PBF_HIP(hipGraphLaunch(S.graph, S.stream));
}

void need(bool c, const char *m) {
  if (!c) invalid(m);
}

struct Host {
  std::vector<float4> pos, vel, x;
  std::vector<int> pid, ids, key, nbr;
  std::vector<float> a, b;
};

}  // namespace

extern "C" {

int pbf_params_init(pbf_params *p, int32_t physics) {
  return guard([&] {
    need(p != nullptr, "params is NULL");
    if (physics != PBF_PHYSICS_REF && physics != PBF_PHYSICS_PAPER)
      invalid("unknown physics mode");
    std::memset(p, 0, sizeof *p);
    p->physics = physics;
    p->iterations = 4;           // Simulation.cpp:503
    p->gravity_y = -9.81f;       // predict_positions.cl:8
    p->xsph_c = 0.01f;           // apply_vorticity_and_viscosity.cl:103
    p->precision = PBF_PRECISION_FP32;
    p->fp32_iterations = 1;
    if (physics == PBF_PHYSICS_REF) {
      p->relaxation = 10000.0f;  // compute_scaling.cl:16
    } else {
      p->relaxation = 600.0f;
      // the stable decade at rho0 = 6378, h = 0.1, m = 1, eps = 600; k = 0.1
      // diverges there (the artificial pressure at rest spacing is ~200x
      // lambda)
      p->scorr_k = 0.001f;
      p->scorr_n = 4.0f;
      p->scorr_dq = 0.2f;
      p->vorticity_eps = 1e-4f;
    }
  });
}

int pbf_create(const pbf_params *params, const pbf_particle *particles,
               uint32_t n, pbf_sim **out) {
  return guard([&] {
    need(params != nullptr && out != nullptr, "params/out is NULL");
    *out = nullptr;
    auto *S = new pbf_sim();
    try {
      const pbf_particle *parts[1] = {particles};
      build(*S, 1, params, parts, &n);
    } catch (...) {
      delete S;
      throw;
    }
    *out = S;
  });
}

int pbf_batch_create(uint32_t nscenes, const pbf_params *params,
                     const pbf_particle *const *particles,
                     const uint32_t *counts, pbf_sim **out) {
  return guard([&] {
    need(params && particles && counts && out, "NULL argument");
    *out = nullptr;
    auto *S = new pbf_sim();
    try {
      build(*S, nscenes, params, particles, counts);
    } catch (...) {
      delete S;
      throw;
    }
    *out = S;
  });
}

int pbf_step(pbf_sim *S, uint32_t nsteps) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    if (S->shape.slots == 0 || nsteps == 0) return;
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    ensure_graph(*S);
    S->ensure_predicted();
    launch_steps(*S, nsteps);
  });
}

int pbf_synchronize(pbf_sim *S) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    PBF_HIP(hipStreamSynchronize(S->stream));
    S->check_err();
  });
}

static void dump_scene(pbf_sim &S, uint32_t s, float *pos4, float *vel4) {
  const pbf::Scene &sc = S.scenes[s];
  const size_t n = (size_t)sc.n;
  if (n == 0) return;
  const size_t padded = (n + pbf::kSlotAlign - 1) / pbf::kSlotAlign *
                        pbf::kSlotAlign;
  std::vector<float4> hp(padded), hv(padded);
  std::vector<int> hid(padded);
  PBF_HIP(hipMemcpyAsync(hp.data(), S.d.pos + sc.slot_off,
                         padded * sizeof(float4), hipMemcpyDeviceToHost,
                         S.stream));
  PBF_HIP(hipMemcpyAsync(hv.data(), S.d.vel + sc.slot_off,
                         padded * sizeof(float4), hipMemcpyDeviceToHost,
                         S.stream));
  PBF_HIP(hipMemcpyAsync(hid.data(), S.d.pid + sc.slot_off,
                         padded * sizeof(int), hipMemcpyDeviceToHost,
                         S.stream));
  PBF_HIP(hipStreamSynchronize(S.stream));
  for (size_t k = 0; k < n; ++k) {
    const int id = hid[k] - sc.id_base;
    if (id < 0 || (size_t)id >= n)
      throw ApiError(PBF_ERR_STATE, "corrupt particle id table");
    if (pos4) std::memcpy(pos4 + 4 * (size_t)id, &hp[k], sizeof(float4));
    if (vel4) std::memcpy(vel4 + 4 * (size_t)id, &hv[k], sizeof(float4));
  }
}

int pbf_dump(pbf_sim *S, float *pos4, float *vel4) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    need(S->scenes.size() == 1, "pbf_dump needs a single-scene handle; use "
                                "pbf_batch_dump");
    PBF_HIP(hipSetDevice(S->device));
    dump_scene(*S, 0, pos4, vel4);
    S->check_err();
  });
}

int pbf_batch_dump(pbf_sim *S, uint32_t scene, float *pos4, float *vel4) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    need(scene < S->scenes.size(), "scene index out of range");
    PBF_HIP(hipSetDevice(S->device));
    dump_scene(*S, scene, pos4, vel4);
    S->check_err();
  });
}

int pbf_positions_to_device(pbf_sim *S, uint32_t scene, void *dst) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    need(dst != nullptr, "dst is NULL");
    need(scene < S->scenes.size(), "scene index out of range");
    PBF_HIP(hipSetDevice(S->device));
    // the kernel writes dst: it must be device memory of this handle's GPU
    // (a host pointer would fault the GPU instead of failing here)
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, dst) != hipSuccess) {
      (void)hipGetLastError();
      throw ApiError(PBF_ERR_INVALID, "dst is not memory HIP knows (allocate it "
                                  "with hipMalloc or map a registered "
                                  "graphics buffer)");
    }
    need(a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged,
         "dst is not device memory");
    need(a.device == S->device, "dst is on another GPU than the simulation");
    // dst may have been mapped or filled on any stream (the handle's stream
    // does not wait for the null stream): wait for everything queued on the
    // GPU first, as the reference finishes GL before it acquires the buffer
    // (src/Simulation.cpp:464)
    PBF_HIP(hipDeviceSynchronize());
    pbf::enqueue_store_positions(S->d, S->scenes[scene],
                                 static_cast<float4 *>(dst), S->stream);
    PBF_HIP(hipGetLastError());
    PBF_HIP(hipStreamSynchronize(S->stream));
    S->check_err();
  });
}

int pbf_set_wave_generator(pbf_sim *S, float value) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    S->wave = value;
    for (auto &sc : S->scenes) sc.wave = value;
    S->dirty = true;
  });
}

int pbf_get_box(const pbf_sim *S, float min4[4], float max4[4]) {
  return guard([&] {
    need(S && min4 && max4, "NULL argument");
    const pbf_params &p = S->params[0];
    const float mn[4] = {p.x_min, p.y_min, p.z_min, 0.0f};
    const float mx[4] = {p.x_max, p.y_max, p.z_max, 0.0f};
    std::memcpy(min4, mn, sizeof mn);
    std::memcpy(max4, mx, sizeof mx);
  });
}

uint32_t pbf_num_particles(const pbf_sim *S) {
  return S ? (uint32_t)S->total : 0u;
}

uint32_t pbf_num_scenes(const pbf_sim *S) {
  return S ? (uint32_t)S->scenes.size() : 0u;
}

uint32_t pbf_num_cells(const pbf_sim *S) {
  return S ? (uint32_t)S->shape.cells : 0u;
}

int pbf_destroy(pbf_sim *S) {
  return guard([&] {
    if (S) {
      (void)hipSetDevice(S->device);
      delete S;
    }
  });
}

const char *pbf_last_error(void) { return g_error.c_str(); }

int pbf_probe(pbf_sim *S, int32_t *cell_of, int32_t *sorted_ids,
              int32_t *cell_start, int32_t *nbr_count, float *rho,
              float *lambda, float *pred4) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    need(S->scenes.size() == 1, "pbf_probe needs a single-scene handle");
    const pbf::Scene &sc = S->scenes[0];
    const size_t n = (size_t)sc.n;
    if (n == 0) {
      if (cell_start)
        std::fill(cell_start, cell_start + S->shape.cells + 1, 0);
      return;
    }
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    S->predicted = false;
    pbf::enqueue_probe(S->d, S->shape, S->stream);
    const size_t P = (size_t)S->shape.slots, M = (size_t)S->shape.cells;
    Host h;
    h.pid.resize(P);
    h.key.resize(P);
    h.ids.resize(P);
    h.nbr.resize(P);
    h.a.resize(P);
    h.b.resize(P);
    h.x.resize(P);
    std::vector<int> cs(M + 1);
    auto d2h = [&](void *dst, const void *src, size_t bytes) {
      PBF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                             S->stream));
    };
    d2h(h.pid.data(), S->d.pid, P * sizeof(int));
    d2h(h.key.data(), S->d.key, P * sizeof(int));
    d2h(h.ids.data(), S->d.pid_s, P * sizeof(int));
    d2h(h.nbr.data(), S->d.nbr, P * sizeof(int));
    d2h(h.a.data(), S->d.rho, P * sizeof(float));
    d2h(h.b.data(), S->d.lam, P * sizeof(float));
    d2h(h.x.data(), S->d.xa, P * sizeof(float4));
    d2h(cs.data(), S->d.cell_start, (M + 1) * sizeof(int));
    PBF_HIP(hipStreamSynchronize(S->stream));
    for (size_t k = 0; k < n; ++k) {
      const int cur = h.pid[k] - sc.id_base;    // previous order
      const int srt = h.ids[k] - sc.id_base;    // sorted order
      if (cur < 0 || (size_t)cur >= n || srt < 0 || (size_t)srt >= n)
        throw ApiError(PBF_ERR_STATE, "corrupt particle id table");
      if (cell_of) cell_of[cur] = h.key[k] - sc.cell_base;
      if (sorted_ids) sorted_ids[k] = srt;
      if (nbr_count) nbr_count[srt] = h.nbr[k];
      if (rho) rho[srt] = h.a[k];
      if (lambda) lambda[srt] = h.b[k];
      if (pred4) std::memcpy(pred4 + 4 * (size_t)srt, &h.x[k], sizeof(float4));
    }
    if (cell_start) std::copy(cs.begin(), cs.end(), cell_start);
  });
}

int pbf_diagnostics(pbf_sim *S, double out[8]) {
  return guard([&] {
    need(S && out, "NULL argument");
    std::fill(out, out + 8, 0.0);
    if (S->shape.slots == 0) return;
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    S->predicted = false;
    pbf::enqueue_diag(S->d, S->shape, S->stream, S->diag_out);
    double h[8];
    PBF_HIP(hipMemcpyAsync(h, S->diag_out, sizeof h, hipMemcpyDeviceToHost,
                           S->stream));
    PBF_HIP(hipStreamSynchronize(S->stream));
    // h: max e+, sum |e| (with a neighbour), count (with a neighbour),
    // momentum xyz, energy, max |e| (with a neighbour)
    out[0] = h[0];
    out[1] = h[2] > 0 ? h[1] / h[2] : 0.0;
    out[2] = h[3];
    out[3] = h[4];
    out[4] = h[5];
    out[5] = h[6];
    out[6] = h[2];
    out[7] = h[7];
  });
}

int pbf_profile_steps(pbf_sim *S, uint32_t nsteps, double *ms,
                      uint32_t *launches, const char **names,
                      uint32_t *nclass) {
  return guard([&] {
    need(S && ms && launches && nclass, "NULL argument");
    const uint32_t cap = *nclass;
    *nclass = pbf::kNumClasses;
    for (uint32_t k = 0; k < cap && k < (uint32_t)pbf::kNumClasses; ++k) {
      ms[k] = 0.0;
      launches[k] = 0;
      if (names) names[k] = pbf::kernel_class_name((int)k);
    }
    if (S->shape.slots == 0) return;
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    autotune(*S);
    S->ensure_predicted();
    EventTimer timer;
    for (uint32_t t = 0; t < nsteps; ++t)
      pbf::enqueue_step(S->d, S->shape, S->stream, &timer);
    PBF_HIP(hipStreamSynchronize(S->stream));
    for (const auto &r : timer.recs) {
      float e = 0.0f;
      PBF_HIP(hipEventElapsedTime(&e, r.a, r.b));
      if ((uint32_t)r.cls < cap) {
        ms[r.cls] += e;
        launches[r.cls] += 1;
      }
    }
  });
}

int pbf_time_steps(pbf_sim *S, uint32_t nsteps, double *ms) {
  return guard([&] {
    need(S && ms, "NULL argument");
    *ms = 0.0;
    if (S->shape.slots == 0 || nsteps == 0) return;
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    ensure_graph(*S);
    S->ensure_predicted();
    hipEvent_t a, b;
    PBF_HIP(hipEventCreate(&a));
    PBF_HIP(hipEventCreate(&b));
    PBF_HIP(hipEventRecord(a, S->stream));
    launch_steps(*S, nsteps);
    PBF_HIP(hipEventRecord(b, S->stream));
    PBF_HIP(hipEventSynchronize(b));
    float e = 0.0f;
    PBF_HIP(hipEventElapsedTime(&e, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = e;
  });
}

int pbf_time_kernel(pbf_sim *S, const char *cls, uint32_t reps, double *ms) {
  return guard([&] {
    need(S && cls && ms && reps >= 1, "bad argument");
    *ms = 0.0;
    if (S->shape.slots == 0) return;
    int c = -1;
    for (int k = 0; k < pbf::kNumClasses; ++k)
      if (std::strcmp(cls, pbf::kernel_class_name(k)) == 0) c = k;
    need(c >= 0, "unknown kernel class");
    need(pbf::repeatable_class(S->shape, c),
         "kernel class has no repeatable kernel (search, lambda, delta)");
    PBF_HIP(hipSetDevice(S->device));
    S->upload_scenes();
    ensure_graph(*S);
    S->ensure_predicted();
    PBF_HIP(hipGraphLaunch(S->graph, S->stream));  // fresh temporaries
    // the launches are captured into a graph, as in a step
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    auto cleanup = [&] {
      if (ge) (void)hipGraphExecDestroy(ge);
      if (g) (void)hipGraphDestroy(g);
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    };
    try {
      PBF_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
      pbf::enqueue_repeat(S->d, S->shape, S->stream, c, (int)reps);
      PBF_HIP(hipStreamEndCapture(S->stream, &g));
      PBF_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      PBF_HIP(hipGraphLaunch(ge, S->stream));  // warm-up
      PBF_HIP(hipEventCreate(&a));
      PBF_HIP(hipEventCreate(&b));
      PBF_HIP(hipEventRecord(a, S->stream));
      PBF_HIP(hipGraphLaunch(ge, S->stream));
      PBF_HIP(hipEventRecord(b, S->stream));
      PBF_HIP(hipEventSynchronize(b));
      float e = 0.0f;
      PBF_HIP(hipEventElapsedTime(&e, a, b));
      *ms = e / reps;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int pbf_step_noalloc(const pbf_params *params, float *pos4, float *vel4,
                     uint32_t n, uint32_t nsteps, void *stream) {
  return guard([&] {
    need(params != nullptr, "params is NULL");
    validate(*params);
    if (n == 0 || nsteps == 0) return;
    need(pos4 && vel4, "pos4/vel4 is NULL");
    pbf::Scene sc = make_scene(*params);
    sc.n = (int)n;  // one scene from slot 0: offsets, shift, bases 0
    const long long slots =
        ((long long)n + pbf::kSlotAlign - 1) / pbf::kSlotAlign *
        pbf::kSlotAlign;
    if (slots > (1LL << 30)) invalid("problem too large (slots beyond 2^30)");
    sc.slots_all = (int)slots;
    pbf::StepShape sh = make_shape(*params, &sc, 1, params, slots, sc.ncells);
    // the buffers of the step, 256-byte aligned, from the start of the
    // scratch: first their extent, then their places
    size_t need_bytes = 0;
    auto bump = [&](size_t base) {
      return [&need_bytes, base](size_t bytes) {
        const size_t at = (need_bytes + 255) / 256 * 256;
        need_bytes = at + bytes;
        return reinterpret_cast<void *>(base + at);
      };
    };
    pbf::DevBuffers d{};
    layout_buffers(d, sh, 1, bump(0));
    if (need_bytes > pbf::kStaticScratchBytes) {
      char msg[256];
      std::snprintf(msg, sizeof msg,
                    "pbf_step_noalloc: %u particles need %zu bytes of device "
                    "scratch; the static scratch holds %zu bytes; use "
                    "pbf_create",
                    n, need_bytes, pbf::kStaticScratchBytes);
      invalid(msg);
    }
    int dev = 0;
    PBF_HIP(hipGetDevice(&dev));
    PBF_HIP(hipDeviceGetAttribute(
        &sh.cus, hipDeviceAttributeMultiprocessorCount, dev));
    void *base = nullptr;
    PBF_HIP(pbf::static_scratch(&base));
    need_bytes = 0;
    layout_buffers(d, sh, 1, bump(reinterpret_cast<size_t>(base)));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    pbf::enqueue_static_setup(d, sh, sc, st);
    pbf::enqueue_load_state(d, sh, reinterpret_cast<const float4 *>(pos4),
                            reinterpret_cast<const float4 *>(vel4), (int)n,
                            st);
    pbf::enqueue_predict_zeroed(d, sh, st);
    for (uint32_t t = 0; t < nsteps; ++t)
      pbf::enqueue_step(d, sh, st, nullptr);
    pbf::enqueue_store_state(d, sh, reinterpret_cast<float4 *>(pos4),
                             reinterpret_cast<float4 *>(vel4), (int)n, st);
    PBF_HIP(hipGetLastError());
  });
}

int pbf_calibrate_copy(uint64_t bytes, uint32_t reps, double *ms) {
  return guard([&] {
    need(ms != nullptr && bytes >= 16 && reps >= 1, "bad argument");
    const size_t n = (size_t)(bytes / sizeof(float4));
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (st) (void)hipStreamDestroy(st);
    };
    try {
      PBF_HIP(hipMalloc(&a, n * sizeof(float4)));
      PBF_HIP(hipMalloc(&b, n * sizeof(float4)));
      PBF_HIP(hipMemset(a, 0, n * sizeof(float4)));
      PBF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      PBF_HIP(hipEventCreate(&e0));
      PBF_HIP(hipEventCreate(&e1));
      pbf::enqueue_copy(a, b, n, st);  // warm-up
      PBF_HIP(hipEventRecord(e0, st));
      for (uint32_t r = 0; r < reps; ++r) pbf::enqueue_copy(a, b, n, st);
      PBF_HIP(hipEventRecord(e1, st));
      PBF_HIP(hipEventSynchronize(e1));
      float t = 0.0f;
      PBF_HIP(hipEventElapsedTime(&t, e0, e1));
      *ms = t / reps;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int pbf_get_launch(const pbf_sim *S, int32_t *block, int32_t *grid,
                   int32_t n, int32_t *count, int32_t *scan_cells) {
  return guard([&] {
    need(S != nullptr, "sim is NULL");
    need(n >= 0 && (n == 0 || (block && grid)), "bad argument");
    for (int c = 0; c < n && c < pbf::kNumClasses; ++c) {
      block[c] = S->shape.block[c];
      grid[c] = S->shape.grid[c];
    }
    if (count) *count = pbf::kNumClasses;
    if (scan_cells) *scan_cells = pbf::kScanTile * S->shape.scan_v;
  });
}

int pbf_get_config(const pbf_sim *S, int32_t cfg[4]) {
  return guard([&] {
    need(S && cfg, "NULL argument");
    cfg[0] = S->shape.block[S->shape.physics == pbf::kPaper
                                ? pbf::kClsSearch
                                : pbf::kClsLambda];
    cfg[1] = 1;
    cfg[2] = 1;
    cfg[3] = S->shape.fp16_iterations;
  });
}

}  // extern "C"
