// simulation.hpp — C++ host facade with the reference's Simulation API
// (matter/src/simulation/simulation.hpp:25-281) over the C-ABI of libmpmx.
//
// It keeps the lifecycle the renderer adapter drives
// (raytracing/src/mpm/MPMIntegrationSim.cpp:16-116):
//   initializeBasic(name) -> setupScene(...) -> prepareSimulation() ->
//   loop { step(); if (frameFinished()) read particles; } until isFinished()
// and the public state it reads (Np, dx, particles.x / pigments /
// diffusion_factor / active, getGridBoundaries, getSpatulaObject).  Particle
// arrays are a host mirror refreshed by syncParticles() (one device->host copy);
// the device state is authoritative.  Header-only; link with libmpmx.so.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mpmx.h"

namespace matter_amd {

typedef float T;
typedef std::array<float, 3> TV;
typedef std::array<float, 9> TM;          // row-major 3x3
typedef std::array<float, 7> Latent;

enum class BC { NoSlip = MPMX_BC_NOSLIP, SlipStick = MPMX_BC_SLIPSTICK, SlipFree = MPMX_BC_SLIPFREE };
enum class PlateType { top = 0, bottom = 1, left = 2, right = 3, front = 4, back = 5 };
enum class PlasticModel { NoPlasticity = MPMX_PLASTIC_NONE, VMVisc = MPMX_PLASTIC_VMVISC, DPVisc = MPMX_PLASTIC_DPVISC };

// Particles (data_structures.hpp:9-50): the fields the path and the renderer use.
struct Particles {
  std::vector<uint8_t> active;
  std::vector<TV> x, v;
  std::vector<TM> F, Bmat;
  std::vector<T> eps_pl_dev, delta_gamma, diffusion_factor;
  std::vector<T> eps_pl_vol, eps_pl_vol_pradhana;
  std::vector<Latent> pigments;
  void resize(size_t n) {
    active.assign(n, 1);
    x.assign(n, TV{0, 0, 0});
    v.assign(n, TV{0, 0, 0});
    F.assign(n, TM{1, 0, 0, 0, 1, 0, 0, 0, 1});
    Bmat.assign(n, TM{});
    eps_pl_dev.assign(n, 0);
    eps_pl_vol.assign(n, 0);
    eps_pl_vol_pradhana.assign(n, 0);
    delta_gamma.assign(n, 0);
    diffusion_factor.assign(n, 0);
    pigments.assign(n, Latent{});
  }
};

// ObjectPlate (object_plate.hpp:13-78)
struct ObjectPlate {
  T pos_object;
  PlateType plate_type;
  BC bc;
  T friction;
  T pos_lower = -1e15f, pos_upper = 1e15f;
  T vx_object = 0, vy_object = 0, vz_object = 0;
  T vmin_factor = 1, load_factor = 0;   // object_plate.hpp:13
};

// Knife (ObjectSpatula's role, object_spatula.hpp): shape, BC and motion.
struct ObjectKnife {
  mpmx_knife spec;
  mpmx_knife_state state;   // state of the current step (invTransform etc.)
};

class Simulation {
 public:
  Simulation() { mpmx_config_default(&cfg); }
  ~Simulation() {
    if (h_) mpmx_destroy(h_);
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;

  // ---- public members (simulation.hpp:30-126)
  int exit = 0;
  unsigned int end_frame = 10;
  unsigned int frame = 0;
  bool is_initialized = false;
  unsigned int Np = 0;
  T dx = 0.01f;
  T particle_mass = 0, particle_volume = 0;
  mpmx_config cfg;
  Particles particles;
  std::vector<ObjectPlate> plates;
  std::unique_ptr<ObjectKnife> knife;
  std::string sim_name;
  std::string directory = "output/";   // simulation.hpp:205

  // ---- lifecycle (simulation_units.cpp:9-245)
  void initializeBasic(const std::string& name) {
    sim_name = name;
    is_initialized = true;
  }

  // Load the two JSON files the reference reads (simulation_units.cpp:32-80).
  int loadConfig(const std::string& sim_config_json, const std::string& pigment_config_json) {
    return mpmx_config_load_json(&cfg, sim_config_json.c_str(), pigment_config_json.c_str());
  }

  // Paint scene: box blobs (one per latent) side by side along x, the floor
  // (SlipStick, mu = 1; simulation_units.cpp:117) and a knife.
  void setupScene(float fps_value, const std::vector<mpmx_box>& blobs, const ObjectKnife* knife_in) {
    cfg.fps = fps_value;
    cfg.dx = dx;
    int64_t n = 0;
    check(mpmx_scene_boxes(blobs.data(), (int)blobs.size(), dx, 2, 42, nullptr, nullptr, 0, &n), "scene");
    std::vector<float> x(3 * n), pig(7 * n);
    check(mpmx_scene_boxes(blobs.data(), (int)blobs.size(), dx, 2, 42, x.data(), pig.data(), n, &n), "scene");
    Np = (unsigned)n;
    particles.resize(Np);
    for (unsigned p = 0; p < Np; ++p) {
      particles.x[p] = TV{x[3 * p], x[3 * p + 1], x[3 * p + 2]};
      for (int c = 0; c < 7; ++c) particles.pigments[p][c] = pig[7 * p + c];
    }
    plates.clear();
    plates.push_back(ObjectPlate{0.0f, PlateType::bottom, BC::SlipStick, 1.0f});
    if (knife_in) knife.reset(new ObjectKnife(*knife_in));
  }

  // Simulation::prepareSimulation (simulation_units.cpp:133-215): derived
  // constants, lattice, first scatter — here on the device.
  void prepareSimulation() {
    if (!is_initialized) throw std::runtime_error("Simulation not initialized");
    cfg.dx = dx;
    if (h_) mpmx_destroy(h_);
    h_ = nullptr;
    check(mpmx_create(&cfg, &h_), "create");
    upload();
    std::vector<mpmx_plate> pl;
    for (const ObjectPlate& p : plates) {
      mpmx_plate q;
      q.pos_object = p.pos_object;
      q.plate_type = (int)p.plate_type;
      q.bc = (int)p.bc;
      q.friction = p.friction;
      q.pos_lower = p.pos_lower;
      q.pos_upper = p.pos_upper;
      q.v[0] = p.vx_object; q.v[1] = p.vy_object; q.v[2] = p.vz_object;
      q.vmin_factor = p.vmin_factor;
      q.load_factor = p.load_factor;
      pl.push_back(q);
    }
    check(mpmx_set_objects(h_, pl.data(), (int)pl.size(), knife ? &knife->spec : nullptr, knife ? 1 : 0),
          "set_objects");
    check(mpmx_prepare(h_), "prepare");
    particle_volume = dx * dx * dx / cfg.ppc;
    particle_mass = cfg.rho * particle_volume;
    int dims[3];
    float org[3];
    check(mpmx_get_lattice(h_, dims, org, nullptr, nullptr, nullptr), "lattice");
    for (int d = 0; d < 3; ++d) {
      low_[d] = org[d];
      high_[d] = org[d] + (dims[d] - 1) * dx;
    }
    frame = 0;
    time_ = 0;
  }

  // Simulation::step (simulation_units.cpp:217-233): one advanceStep on the device.
  void step() {
    const int rc = mpmx_step(h_, 1);
    if (rc == MPMX_ERR_DEVICE) exit = 1;  // reference exit flag
    else check(rc, "step");
    time_ += cfg.dt;
    if (knife) mpmx_knife_state_at(h_, 0, stepsDone(), &knife->state);
  }

  bool frameFinished() {
    int f = 0;
    check(mpmx_frame_finished(h_, &f), "frame_finished");
    if (f) frame++;
    return f != 0;
  }
  bool isFinished() const { return frame >= end_frame; }

  // Simulation::saveParticleData / saveGridData (save_data.cpp:7-238): PLY
  // files at the reference's paths, directory + sim_name + "/particles_f" +
  // extra + frame + ".ply" (and "/grid_f...").
  void saveParticleData(const std::string& extra = "") {
    check(mpmx_write_particles_ply(h_, (directory + sim_name + "/particles_f" + extra + std::to_string(frame) + ".ply").c_str()),
          "saveParticleData");
  }
  void saveGridData(const std::string& extra = "") {
    check(mpmx_write_grid_ply(h_, (directory + sim_name + "/grid_f" + extra + std::to_string(frame) + ".ply").c_str()),
          "saveGridData");
  }

  // Refresh the host mirror of the particles (what the adapter reads).
  void syncParticles() {
    std::vector<float> x(3 * Np), v(3 * Np), F(9 * Np), C(9 * Np), pig(7 * Np), diff(Np), eps(Np), dg(Np);
    std::vector<uint8_t> act(Np);
    check(mpmx_download_particles(h_, x.data(), v.data(), F.data(), C.data(), pig.data(), diff.data(), eps.data(),
                                  dg.data(), act.data()),
          "download");
    check(mpmx_download_plastic_state(h_, particles.eps_pl_vol.data(), particles.eps_pl_vol_pradhana.data()),
          "download");
    for (unsigned p = 0; p < Np; ++p) {
      for (int d = 0; d < 3; ++d) {
        particles.x[p][d] = x[3 * p + d];
        particles.v[p][d] = v[3 * p + d];
      }
      for (int c = 0; c < 9; ++c) {
        particles.F[p][c] = F[9 * p + c];
        particles.Bmat[p][c] = C[9 * p + c];
      }
      for (int c = 0; c < 7; ++c) particles.pigments[p][c] = pig[7 * p + c];
      particles.diffusion_factor[p] = diff[p];
      particles.eps_pl_dev[p] = eps[p];
      particles.delta_gamma[p] = dg[p];
      particles.active[p] = act[p];
    }
  }

  // Renderer format straight from the device (MPMIntegrationSim::recountParticles).
  void packRender(std::vector<std::array<float, 4>>& pos4, std::vector<Latent>& lat7, std::vector<float>& diff) {
    pos4.resize(Np);
    lat7.resize(Np);
    diff.resize(Np);
    check(mpmx_pack_render(h_, &pos4[0][0], &lat7[0][0], diff.data()), "pack_render");
  }

  std::pair<std::vector<T>, std::vector<T>> getGridBoundaries() const {
    return {std::vector<T>{low_[0], low_[1], low_[2]}, std::vector<T>{high_[0], high_[1], high_[2]}};
  }
  const ObjectKnife* getSpatulaObject() const { return knife.get(); }
  mpmx_handle* handle() const { return h_; }
  double time() const { return time_; }

 private:
  mpmx_handle* h_ = nullptr;
  T low_[3] = {0, 0, 0}, high_[3] = {0, 0, 0};
  double time_ = 0;

  int64_t stepsDone() const {
    mpmx_stats s;
    mpmx_stats_get(h_, &s);
    return s.step;
  }
  void upload() {
    std::vector<float> x(3 * Np), v(3 * Np), F(9 * Np), C(9 * Np), pig(7 * Np), diff(Np), eps(Np);
    for (unsigned p = 0; p < Np; ++p) {
      for (int d = 0; d < 3; ++d) {
        x[3 * p + d] = particles.x[p][d];
        v[3 * p + d] = particles.v[p][d];
      }
      for (int c = 0; c < 9; ++c) {
        F[9 * p + c] = particles.F[p][c];
        C[9 * p + c] = particles.Bmat[p][c];
      }
      for (int c = 0; c < 7; ++c) pig[7 * p + c] = particles.pigments[p][c];
      diff[p] = particles.diffusion_factor[p];
      eps[p] = particles.eps_pl_dev[p];
    }
    check(mpmx_upload_particles(h_, Np, x.data(), v.data(), F.data(), C.data(), pig.data(), diff.data(), eps.data(),
                                particles.active.data()),
          "upload");
  }
  void check(int rc, const char* what) const {
    if (rc != MPMX_OK) {
      std::string m = std::string("mpmx ") + what + ": " + mpmx_status_string(rc);
      if (h_) m += std::string(" — ") + mpmx_last_error(h_);
      throw std::runtime_error(m);
    }
  }
};

}  // namespace matter_amd
