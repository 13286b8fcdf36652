// simulation_api_demo.cpp — drives the C++ Simulation facade the way the
// renderer adapter drives the reference (MPMIntegrationSim.cpp:16-45):
// initializeBasic -> setupScene -> prepareSimulation -> step until a frame
// finishes -> read particles.  Prints one JSON line for the test to check.
#include <cstdio>
#include <cmath>
#include <sys/stat.h>
#include <string>
#include <vector>

#include "../../pigment_mixing_amd/csrc/host/simulation.hpp"

int main(int argc, char** argv) {
  const std::string cfgdir = argc > 1 ? argv[1] : "pigment_mixing_amd/config";
  matter_amd::Simulation sim;
  sim.initializeBasic("api_demo");
  const int rc = sim.loadConfig(cfgdir + "/sim_config.json", cfgdir + "/pigment_config.json");
  // BASELINE material on top of the reference's files
  sim.cfg.E = 1e5f; sim.cfg.nu = 0.3f; sim.cfg.rho = 1500.0f; sim.cfg.flip_ratio = -0.5f;
  sim.cfg.plastic_model = MPMX_PLASTIC_VMVISC; sim.cfg.q_max = sim.cfg.q_min = 50.0f;
  sim.cfg.perzyna_visc = 0.2f; sim.cfg.gravity[1] = -9.81f; sim.dt = 1e-5f;
  sim.cfg.grid_dims[0] = 32; sim.cfg.grid_dims[1] = 16; sim.cfg.grid_dims[2] = 24;
  sim.dx = 3.125e-3f;
  std::vector<mpmx_box> blobs(2);
  const float blue[7] = {0.8644522f, 0.0045213f, 0.0295347f, 0.1014918f, -0.0536479f, -0.0120357f, 0.0034582f};
  const float yell[7] = {0.0039216f, 0.9343413f, 0.0039216f, 0.0578156f, 0.0333659f, 0.0250436f, -0.0572618f};
  for (int b = 0; b < 2; ++b) {
    blobs[b].lo[0] = b * 6 * sim.dx; blobs[b].lo[1] = 0; blobs[b].lo[2] = 0;
    blobs[b].cells[0] = 6; blobs[b].cells[1] = 4; blobs[b].cells[2] = 6;
    for (int c = 0; c < 7; ++c) blobs[b].latent[c] = b ? yell[c] : blue[c];
  }
  sim.setupScene(1000.0f /* fps: a frame every 100 steps */, blobs, nullptr);
  sim.prepareSimulation();
  int steps = 0;
  while (!sim.frameFinished()) {
    sim.step();
    ++steps;
    if (steps > 1000) break;
  }
  sim.syncParticles();
  std::vector<std::array<float, 4>> pos4;
  std::vector<matter_amd::Latent> lat;
  std::vector<float> diff;
  sim.packRender(pos4, lat, diff);
  double ymean = 0;
  float vmin = 0;
  int act = 0;
  for (unsigned p = 0; p < sim.Np; ++p) {
    ymean += sim.particles.x[p][1];
    vmin = std::fmin(vmin, sim.particles.v[p][1]);
    act += sim.particles.active[p];
  }
  // saveParticleData into <outdir>/api_demo/particles_f<frame>.ply
  long ply_bytes = -1;
  if (argc > 2) {
    sim.directory = std::string(argv[2]) + "/";
    mkdir((sim.directory + sim.sim_name).c_str(), 0755);
    sim.saveParticleData();
    struct stat st;
    const std::string f = sim.directory + sim.sim_name + "/particles_f" + std::to_string(sim.frame) + ".ply";
    if (stat(f.c_str(), &st) == 0) ply_bytes = (long)st.st_size;
  }
  const auto gb = sim.getGridBoundaries();
  const unsigned frame1 = sim.frame;
  const double time1 = sim.time();
  // a second frame with the reference's adaptive step (updateDt before every
  // step, dt_max = cfl_elastic dx / sqrt(E / rho) with cfl_elastic 0.1)
  sim.adaptive_dt = true;
  sim.cfl_elastic = 0.1f;
  sim.dt_max = sim.cfl_elastic * sim.dx / std::sqrt(sim.cfg.E / sim.cfg.rho);
  int steps2 = 0;
  float dt_min2 = 1e30f, dt_max2 = 0.0f;
  while (!sim.frameFinished()) {
    sim.step();
    ++steps2;
    dt_min2 = std::fmin(dt_min2, sim.dt);
    dt_max2 = std::fmax(dt_max2, sim.dt);
    if (steps2 > 1000) break;
  }
  std::printf("{\"config_rc\": %d, \"Np\": %u, \"steps\": %d, \"frame\": %u, \"active\": %d, \"ymean\": %.9g, "
              "\"vymin\": %.9g, \"exit\": %d, \"pos_w\": %.3g, \"lo\": [%.6g, %.6g, %.6g], \"hi\": [%.6g, %.6g, %.6g], "
              "\"time\": %.9g, \"ply_bytes\": %ld, \"adaptive\": {\"steps\": %d, \"frame\": %u, \"time\": %.9g, "
              "\"dt_min\": %.9g, \"dt_max\": %.9g, \"dt_bound\": %.9g, \"exit\": %d}}\n",
              rc, sim.Np, steps, frame1, act, ymean / sim.Np, vmin, sim.exit, pos4[0][3], gb.first[0], gb.first[1],
              gb.first[2], gb.second[0], gb.second[1], gb.second[2], time1, ply_bytes, steps2, sim.frame, sim.time(),
              dt_min2, dt_max2, sim.dt_max, sim.exit);
  return 0;
}
