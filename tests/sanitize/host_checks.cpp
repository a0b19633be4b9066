// Host-side checks of the C ABI's scenario readers and parameter validation
// and of the CPU oracle, compiled with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py.  Nothing here
// reaches a GPU: pbf_create is only called with parameters that its
// validation rejects before any device call.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pbf.h"

extern "C" {
int orc_set_threads(int nthreads);
int orc_run_f32(const pbf_params *p, int n, float *pos4, float *vel4,
                int nsteps, float wave);
int orc_run_f64(const pbf_params *p, int n, double *pos4, double *vel4,
                int nsteps, float wave);
int orc_probe_f32(const pbf_params *p, int n, const float *pos4,
                  const float *vel4, int *cell_of, int *sorted_ids,
                  int *cell_start, int *nbr, float *rho, float *lambda,
                  float *pred4);
long orc_num_cells(const pbf_params *p);
}

static int failures = 0;

static void check(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s (%s)\n", what, pbf_last_error());
    ++failures;
  }
}

template <typename T>
static void pack(const std::vector<pbf_particle> &parts, std::vector<T> &pos,
                 std::vector<T> &vel) {
  const size_t n = parts.size();
  pos.assign(4 * n, T(0));
  vel.assign(4 * n, T(0));
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      pos[4 * i + k] = parts[i].x[k];
      vel[4 * i + k] = parts[i].v[k];
    }
    vel[4 * i + 3] = parts[i].m;
  }
}

// Reads a scenario, then runs the oracle on it for a few steps in both
// physics modes (FP32 and FP64) and probes its neighbour search.
static void scenario(const std::string &dir, const char *name) {
  pbf_params p;
  char part_file[512];
  const std::string par = dir + "/" + name + ".par";
  const std::string in = dir + "/" + name + ".in";
  check(pbf_read_par(par.c_str(), &p, part_file, sizeof part_file) == PBF_OK,
        par.c_str());
  uint32_t n = 0;
  check(pbf_read_particles(in.c_str(), nullptr, 0, &n) == PBF_OK && n > 0,
        "count particles");
  std::vector<pbf_particle> parts(n);
  uint32_t got = 0;
  check(pbf_read_particles(in.c_str(), parts.data(), n, &got) == PBF_OK &&
            got == n,
        "read particles");
  if (n > 1) {  // a buffer smaller than the file
    std::vector<pbf_particle> one(1);
    check(pbf_read_particles(in.c_str(), one.data(), 1, &got) == PBF_OK &&
              got == n && std::memcmp(&one[0], &parts[0], sizeof one[0]) == 0,
          "short read");
  }
  std::vector<float> pos, vel;
  pack(parts, pos, vel);
  const long ncell = orc_num_cells(&p);
  check(ncell > 0, "cell count");
  std::vector<int> cell_of(n), sorted(n), start(ncell + 1), nbr(n);
  std::vector<float> rho(n), lam(n), pred(4 * n);
  check(orc_probe_f32(&p, (int)n, pos.data(), vel.data(), cell_of.data(),
                      sorted.data(), start.data(), nbr.data(), rho.data(),
                      lam.data(), pred.data()) == 0,
        "oracle probe");
  check(orc_run_f32(&p, (int)n, pos.data(), vel.data(), 3, 0.0f) == 0,
        "oracle ref f32");
  std::vector<double> pos64, vel64;
  pack(parts, pos64, vel64);
  check(orc_run_f64(&p, (int)n, pos64.data(), vel64.data(), 3, 0.0f) == 0,
        "oracle ref f64");
  pbf_params q = p;
  check(pbf_params_init(&q, PBF_PHYSICS_PAPER) == PBF_OK, "paper params");
  // the scenario's box, cells, density and time step, PAPER physics
  q.time_step_length = p.time_step_length;
  q.x_min = p.x_min, q.x_max = p.x_max;
  q.y_min = p.y_min, q.y_max = p.y_max;
  q.z_min = p.z_min, q.z_max = p.z_max;
  q.x_n = p.x_n, q.y_n = p.y_n, q.z_n = p.z_n;
  q.rest_density = p.rest_density;
  pack(parts, pos64, vel64);
  check(orc_run_f64(&q, (int)n, pos64.data(), vel64.data(), 3, 0.0f) == 0,
        "oracle paper f64");
}

static void write_file(const std::string &path, const char *text) {
  if (FILE *f = std::fopen(path.c_str(), "w")) {
    std::fputs(text, f);
    std::fclose(f);
  }
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::printf("usage: host_checks <golden dir> <scratch dir>\n");
    return 2;
  }
  const std::string golden = argv[1], tmp = argv[2];
  orc_set_threads(2);

  pbf_params p;
  check(pbf_params_init(&p, PBF_PHYSICS_REF) == PBF_OK, "ref params");
  check(pbf_params_init(&p, 7) == PBF_ERR_INVALID, "unknown physics");
  check(std::strlen(pbf_last_error()) > 0, "error text");

  for (const char *name : {"two", "two_updown", "27", "box_small", "box", "dam"})
    scenario(golden, name);

  // malformed and missing inputs
  pbf_params q;
  check(pbf_read_par((tmp + "/missing.par").c_str(), &q, nullptr, 0) != PBF_OK,
        "missing .par");
  write_file(tmp + "/bad.par",
             "x_min not-a-number\nunknown_key 3\npart_input_file\n");
  pbf_read_par((tmp + "/bad.par").c_str(), &q, nullptr, 0);  // no crash
  write_file(tmp + "/short.in", "5\n1 0.1 0.2 0.3 0 0 0\n1 0.4\n");
  uint32_t n = 0;
  std::vector<pbf_particle> parts(5);
  pbf_read_particles((tmp + "/short.in").c_str(), parts.data(), 5, &n);
  check(n <= 5, "short file count");
  write_file(tmp + "/empty.in", "");
  pbf_read_particles((tmp + "/empty.in").c_str(), parts.data(), 5, &n);

  // parameter validation (rejected before any device call)
  pbf_params bad;
  pbf_params_init(&bad, PBF_PHYSICS_PAPER);
  bad.iterations = 0;
  pbf_sim *sim = nullptr;
  check(pbf_create(&bad, parts.data(), 1, &sim) == PBF_ERR_INVALID &&
            sim == nullptr,
        "iterations 0 rejected");
  check(pbf_create(nullptr, parts.data(), 1, &sim) == PBF_ERR_INVALID,
        "NULL params rejected");

  // the entry point that allocates nothing: argument and size checks come
  // before any device call
  float dummy[8] = {0};
  check(pbf_step_noalloc(nullptr, dummy, dummy, 1, 1, nullptr) ==
            PBF_ERR_INVALID,
        "noalloc: NULL params rejected");
  check(pbf_step_noalloc(&bad, dummy, dummy, 1, 1, nullptr) ==
            PBF_ERR_INVALID,
        "noalloc: bad params rejected");
  pbf_params big;
  pbf_params_init(&big, PBF_PHYSICS_PAPER);
  big.x_max = big.y_max = big.z_max = 4.0f;
  big.x_n = big.y_n = big.z_n = 40.0f;
  big.time_step_length = 0.01f;
  big.rest_density = 6378.0f;
  check(pbf_step_noalloc(&big, dummy, dummy, 0, 1, nullptr) == PBF_OK,
        "noalloc: no particles is a no-op");
  check(pbf_step_noalloc(&big, nullptr, dummy, 5, 1, nullptr) ==
            PBF_ERR_INVALID,
        "noalloc: NULL state rejected");
  check(pbf_step_noalloc(&big, dummy, dummy, 1000000, 1, nullptr) ==
                PBF_ERR_INVALID &&
            std::strstr(pbf_last_error(), "scratch") != nullptr,
        "noalloc: beyond the static scratch rejected");

  // the viewer's positions buffer: a NULL handle is refused before any
  // device call
  check(pbf_positions_to_device(nullptr, 0, dummy) == PBF_ERR_INVALID,
        "positions_to_device: NULL sim rejected");

  std::printf("sanitized host checks: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}
